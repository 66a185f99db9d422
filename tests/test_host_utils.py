"""Host utilities of the texture and mesh paths: the PNG reader against
PIL on the reference's texture and on generated images of every colour
type; the OBJ loader against a direct parse of the teapot."""
import ctypes
import os

import numpy as np
import pytest

from scene_runner import REPO, RES_DIR

LIB = os.path.join(REPO, "srp_amd", "libsrp_amd.so")


@pytest.fixture(scope="module")
def lib():
    l = ctypes.CDLL(LIB, mode=ctypes.RTLD_LOCAL)
    l.srpPngLoadRGB.restype = ctypes.POINTER(ctypes.c_uint8)
    l.srpPngLoadRGB.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_char_p)]
    return l


def load_png(lib, path):
    w, h, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_char_p()
    p = lib.srpPngLoadRGB(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(err))
    if not p:
        return None, err.value.decode()
    arr = np.ctypeslib.as_array(p, shape=(h.value, w.value, 3)).copy()
    ctypes.CDLL(None).free(p)
    return arr, None


def test_png_stonewall(lib):
    from PIL import Image
    got, err = load_png(lib, os.path.join(RES_DIR, "stoneWall.png"))
    assert err is None
    want = np.asarray(Image.open(os.path.join(RES_DIR, "stoneWall.png")).convert("RGB"))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "1", "I;16"])
def test_png_modes(lib, tmp_path, mode):
    from PIL import Image
    rng = np.random.default_rng(7)
    rgb = rng.integers(0, 256, size=(37, 53, 3), dtype=np.uint8)
    img = Image.fromarray(rgb, "RGB")
    if mode == "P":
        img = img.quantize(colors=200)
    elif mode == "I;16":
        img = Image.fromarray((rng.integers(0, 65536, size=(37, 53))).astype(np.uint16))
    else:
        img = img.convert(mode)
    path = str(tmp_path / f"img_{mode.replace(';', '_')}.png")
    img.save(path)
    got, err = load_png(lib, path)
    assert err is None, err
    ref = Image.open(path)
    if mode == "I;16":
        want16 = np.asarray(ref).astype(np.uint16)
        want = np.repeat((want16 >> 8).astype(np.uint8)[..., None], 3, axis=-1)
    else:
        want = np.asarray(ref.convert("RGB"))
    assert np.array_equal(got, want)


def test_png_missing(lib):
    got, err = load_png(lib, "/nonexistent/file.png")
    assert got is None and err == "can't fopen"


def test_obj_teapot(lib):
    class Mesh(ctypes.Structure):
        _fields_ = [("vertices", ctypes.POINTER(ctypes.c_float)), ("vertexCount", ctypes.c_size_t),
                    ("indices", ctypes.POINTER(ctypes.c_uint32)), ("indexCount", ctypes.c_size_t)]
    m = Mesh()
    path = os.path.join(RES_DIR, "utah_teapot.obj")
    assert lib.srpObjLoad(path.encode(), ctypes.byref(m))
    verts = np.ctypeslib.as_array(m.vertices, shape=(m.vertexCount, 8)).copy()
    idx = np.ctypeslib.as_array(m.indices, shape=(m.indexCount,)).copy()
    lib.srpObjFree(ctypes.byref(m))
    v, vt, vn, faces = [], [], [], []
    for line in open(path):
        if line.startswith("v "):
            v.append([np.float32(x) for x in line.split()[1:4]])
        elif line.startswith("vt"):
            vt.append([np.float32(x) for x in line.split()[1:3]])
        elif line.startswith("vn"):
            vn.append([np.float32(x) for x in line.split()[1:4]])
        elif line.startswith("f"):
            faces.append([tuple(int(t) for t in c.split("/")) for c in line.split()[1:4]])
    want = np.array([v[a - 1] + vt[b - 1] + vn[c - 1] for f in faces for (a, b, c) in f], np.float32)
    assert verts.shape == want.shape == (3498, 8)
    assert np.array_equal(verts, want)
    assert np.array_equal(idx, np.arange(3498, dtype=np.uint32))


def test_obj_bad_face_is_skipped_whole(lib, tmp_path):
    """A face with an out-of-range second or third corner is skipped as a
    whole (nothing of it pushed), so the faces after it stay aligned."""
    class Mesh(ctypes.Structure):
        _fields_ = [("vertices", ctypes.POINTER(ctypes.c_float)), ("vertexCount", ctypes.c_size_t),
                    ("indices", ctypes.POINTER(ctypes.c_uint32)), ("indexCount", ctypes.c_size_t)]
    lines = ["v 0 0 0", "v 1 0 0", "v 0 1 0", "v 1 1 0", "vt 0 0", "vt 1 1", "vn 0 0 1",
             "f 1/1/1 2/2/1 3/1/1",
             "f 1/1/1 9/2/1 3/1/1",      # second corner out of range
             "f 2/1/1 3/2/1 7/1/1",      # third corner out of range
             "f 2/1/1 4/2/1 3/1/1"]
    path = tmp_path / "bad.obj"
    path.write_text("\n".join(lines) + "\n")
    m = Mesh()
    assert lib.srpObjLoad(str(path).encode(), ctypes.byref(m))
    verts = np.ctypeslib.as_array(m.vertices, shape=(m.vertexCount, 8)).copy()
    idx = np.ctypeslib.as_array(m.indices, shape=(m.indexCount,)).copy()
    lib.srpObjFree(ctypes.byref(m))
    assert m.indexCount == 6 and idx.tolist() == list(range(6))
    assert verts[:, :3].tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]


def _png_chunk(kind, data):
    import struct
    import zlib
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


@pytest.mark.parametrize("w,h", [(1 << 31, 1 << 31), (1 << 25, 1), (1, (1 << 24) + 1), (0x40000000, 4)])
def test_png_oversized_header_is_refused(lib, tmp_path, w, h):
    """An IHDR whose sizes would overflow the decoder's buffer arithmetic
    (the row length times the height wraps in 32 bits for 2^31 x 2^31
    RGBA8) is refused before any allocation, as stb_image refuses sides
    above 2^24, with a small valid IDAT behind it."""
    import struct
    import zlib
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    png = b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(b"\x00" * 5)) + \
        _png_chunk(b"IEND", b"")
    path = tmp_path / "huge.png"
    path.write_bytes(png)
    got, err = load_png(lib, str(path))
    assert got is None and err == "image too large"
