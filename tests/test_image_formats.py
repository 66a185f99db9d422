"""Texture files other than PNG (srp_amd/csrc/host/image_decode.c): BMP, TGA
and binary PNM decode to the same RGB bytes as the reference's loader.

The reference reads textures with its vendored stb_image v2.30
(core/texture.c:34, 3 channels).  The images here are written by PIL where
it can (24-bit and palette BMPs, raw and run-length TGAs, PPM/PGM) and by
hand where it cannot (16-bit and bit-field BMPs, OS/2 and top-down BMPs,
15/16-bit and colour-mapped TGAs).  Each is decoded by the decoder built
here and compared with:
  - stbi_load of the reference compiled in place (oracle/_ref), when the
    reference sources are present: every byte must agree;
  - PIL's own decoding (always): every byte must agree, except for the
    16-bit colour formats, where PIL widens 5- and 6-bit channels
    differently and stb's bytes are the ones to match.
JPEG and GIF are refused with a message (stb would decode them: a
documented difference)."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from scene_runner import REPO

HOST = os.path.join(REPO, "srp_amd", "csrc", "host")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libscenes_ref.so")


@pytest.fixture(scope="module")
def dec(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    so = str(tmp_path_factory.mktemp("imgdec") / "libimgdec.so")
    subprocess.run([cc, "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Werror",
                    os.path.join(HOST, "image_decode.c"), os.path.join(HOST, "png_decode.c"), "-lz", "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.srpImageLoadRGB.restype = ctypes.c_void_p
    lib.srpImageLoadRGB.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_char_p)]
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]

    def load(path):
        w, h, err = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_char_p()
        p = lib.srpImageLoadRGB(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(err))
        if not p:
            return None, err.value.decode()
        n = w.value * h.value * 3
        out = np.frombuffer(ctypes.string_at(p, n), np.uint8).reshape(h.value, w.value, 3).copy()
        libc.free(p)
        return out, None
    return load


@pytest.fixture(scope="module")
def stb():
    """stbi_load of the reference compiled in place, or None."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO, mode=ctypes.RTLD_LOCAL)
    lib.stbi_load.restype = ctypes.c_void_p
    lib.stbi_load.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.c_int]
    lib.stbi_image_free.argtypes = [ctypes.c_void_p]
    lib.stbi_failure_reason.restype = ctypes.c_char_p

    def load(path):
        w, h, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        p = lib.stbi_load(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c), 3)
        if not p:
            return None, lib.stbi_failure_reason().decode()
        n = w.value * h.value * 3
        out = np.frombuffer(ctypes.string_at(p, n), np.uint8).reshape(h.value, w.value, 3).copy()
        lib.stbi_image_free(p)
        return out, None
    return load


def _check(dec, stb, path, pil_rgb=None):
    ours, err = dec(path)
    assert err is None, f"{os.path.basename(path)}: {err}"
    if pil_rgb is not None:
        assert ours.shape == pil_rgb.shape and np.array_equal(ours, pil_rgb), os.path.basename(path)
    if stb is not None:
        ref, rerr = stb(path)
        assert rerr is None, rerr
        assert ours.shape == ref.shape and np.array_equal(ours, ref), f"{os.path.basename(path)} differs from stb"
    return ours


def _pil_rgb(img):
    return np.asarray(img.convert("RGB"))


SIZES = [(1, 1), (3, 2), (17, 5), (64, 33), (129, 7)]


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("mode", ["RGB", "P", "L", "1"])
def test_bmp_written_by_pil(dec, stb, tmp_path, w, h, mode):
    from PIL import Image
    rng = np.random.default_rng(w * 131 + h)
    rgb = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    img = Image.fromarray(rgb, "RGB")
    if mode == "P":
        img = img.quantize(colors=200)
    elif mode != "RGB":
        img = img.convert(mode)
    path = str(tmp_path / f"t_{mode}.bmp")
    img.save(path)
    _check(dec, stb, path, _pil_rgb(Image.open(path)))


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("mode,rle", [("RGB", False), ("RGB", True), ("RGBA", False), ("RGBA", True),
                                      ("L", False), ("L", True), ("P", False)])
def test_tga_written_by_pil(dec, stb, tmp_path, w, h, mode, rle):
    from PIL import Image
    rng = np.random.default_rng(w * 7 + h)
    arr = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    arr[: h // 2, : w // 2] = arr[0, 0]   # runs for the encoder
    img = Image.fromarray(arr, "RGBA")
    img = img.quantize(colors=100) if mode == "P" else img.convert(mode)
    path = str(tmp_path / f"t_{mode}_{int(rle)}.tga")
    img.save(path, compression="tga_rle" if rle else None)
    _check(dec, stb, path, _pil_rgb(Image.open(path)))


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("mode", ["RGB", "L"])
def test_pnm_written_by_pil(dec, stb, tmp_path, w, h, mode):
    from PIL import Image
    rng = np.random.default_rng(w + 1000 * h)
    img = Image.fromarray(rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8), "RGB").convert(mode)
    path = str(tmp_path / ("t.ppm" if mode == "RGB" else "t.pgm"))
    img.save(path)
    _check(dec, stb, path, _pil_rgb(Image.open(path)))


def _bmp(w, h, bpp, pixel_rows, hsz=40, compress=0, masks=None, palette=None, top_down=False):
    """A BMP file: rows given bottom-up as raw bytes (padding added here)."""
    pal = b""
    if palette is not None:
        pal = b"".join(bytes([b, g, r]) + (b"" if hsz == 12 else b"\0") for (r, g, b) in palette)
    extra = b""
    if hsz == 40 and compress == 3:
        extra = struct.pack("<3I", *masks)
    if hsz == 12:
        info = struct.pack("<IHHHH", 12, w, h, 1, bpp)
    else:
        info = struct.pack("<IiiHHIIiiII", hsz, w, -h if top_down else h, 1, bpp, compress, 0, 2835, 2835, 0, 0)
        if hsz == 108:
            m = masks or (0, 0, 0, 0)
            info += struct.pack("<4I", *m) + b"\0" * (108 - 40 - 16)
    rows = b""
    for row in pixel_rows:
        rows += row + b"\0" * ((-len(row)) % 4)
    off = 14 + len(info) + len(extra) + len(pal)
    return b"BM" + struct.pack("<IHHI", off + len(rows), 0, 0, off) + info + extra + pal + rows


def test_bmp_16_bit_and_bit_fields_match_stb(dec, stb, tmp_path):
    """555 (default masks), 565 and odd masks (BI_BITFIELDS), 32-bit with
    custom masks, a V4 header, an OS/2 header, a top-down file."""
    if stb is None:
        pytest.skip("reference not compiled in place (oracle/_ref): only stb fixes these texels")
    rng = np.random.default_rng(5)
    w, h = 13, 6
    files = {}
    v16 = rng.integers(0, 1 << 16, size=(h, w), dtype=np.uint16)
    rows16 = [v16[y].astype("<u2").tobytes() for y in range(h)]
    files["555.bmp"] = _bmp(w, h, 16, rows16)
    files["565.bmp"] = _bmp(w, h, 16, rows16, compress=3, masks=(0xF800, 0x07E0, 0x001F))
    files["odd16.bmp"] = _bmp(w, h, 16, rows16, compress=3, masks=(0x7000, 0x0E00, 0x01C0))
    v32 = rng.integers(0, 1 << 32, size=(h, w), dtype=np.uint32)
    rows32 = [v32[y].astype("<u4").tobytes() for y in range(h)]
    files["x32.bmp"] = _bmp(w, h, 32, rows32, compress=3, masks=(0x0FF00000, 0x000FF000, 0x00000FF0))
    wide = str(tmp_path / "wide_masks.bmp")
    with open(wide, "wb") as f:   # channels wider than 8 bits: refused by both
        f.write(_bmp(w, h, 32, rows32, compress=3, masks=(0x3FF00000, 0x000FFC00, 0x000003FF)))
    assert dec(wide) == (None, "bad masks") and stb(wide)[1] == "bad masks"
    files["v4_32.bmp"] = _bmp(w, h, 32, rows32, hsz=108, compress=3, masks=(0xFF, 0xFF00, 0xFF0000, 0xFF000000))
    rows24 = [rng.integers(0, 256, size=3 * w, dtype=np.uint8).tobytes() for _ in range(h)]
    files["os2_24.bmp"] = _bmp(w, h, 24, rows24, hsz=12)
    files["topdown.bmp"] = _bmp(w, h, 24, rows24, top_down=True)
    pal = [tuple(int(x) for x in rng.integers(0, 256, 3)) for _ in range(16)]
    rows4 = [bytes(((int(a) << 4) | int(b)) for a, b in rng.integers(0, 16, size=(7, 2))) for _ in range(h)]
    files["pal4.bmp"] = _bmp(w, h, 4, rows4, palette=pal)
    # an OS/2 header: the reference's loader takes the palette size as
    # (offset - 38) / 3 entries, 12 of these 16; indices past it read
    # uninitialised memory there, so they are not used
    files["os2_pal8.bmp"] = _bmp(w, h, 8, [rng.integers(0, 12, size=w, dtype=np.uint8).tobytes() for _ in range(h)],
                                 hsz=12, palette=pal)
    for name, data in files.items():
        p = str(tmp_path / name)
        with open(p, "wb") as f:
            f.write(data)
        _check(dec, stb, p)


def _tga(w, h, itype, bpp, pixels, cmap=None, cmap_bits=0, top=False, id_bytes=b""):
    hdr = struct.pack("<BBBHHBHHHHBB", len(id_bytes), 1 if cmap is not None else 0, itype, 0,
                      0 if cmap is None else len(cmap) // (cmap_bits // 8 if cmap_bits != 15 else 2),
                      cmap_bits, 0, 0, w, h, bpp, 0x20 if top else 0)
    return hdr + id_bytes + (cmap or b"") + pixels


def test_tga_16_bit_colour_mapped_and_rle_match_stb(dec, stb, tmp_path):
    if stb is None:
        pytest.skip("reference not compiled in place (oracle/_ref): only stb fixes these texels")
    rng = np.random.default_rng(9)
    w, h = 11, 7
    files = {}
    px16 = rng.integers(0, 1 << 16, size=w * h, dtype=np.uint16).astype("<u2").tobytes()
    files["rgb16.tga"] = _tga(w, h, 2, 16, px16)
    files["rgb15_top.tga"] = _tga(w, h, 2, 15, px16, top=True, id_bytes=b"texture")
    files["grey16.tga"] = _tga(w, h, 3, 16, px16)
    cmap24 = rng.integers(0, 256, size=40 * 3, dtype=np.uint8).tobytes()
    idx = rng.integers(0, 45, size=w * h, dtype=np.uint8).tobytes()   # some indices past the map
    files["cmap24.tga"] = _tga(w, h, 1, 8, idx, cmap=cmap24, cmap_bits=24)
    cmap16 = rng.integers(0, 1 << 16, size=40, dtype=np.uint16).astype("<u2").tobytes()
    files["cmap16.tga"] = _tga(w, h, 1, 8, idx, cmap=cmap16, cmap_bits=16)
    # RLE with packets that run across rows, raw packets, 32-bit pixels
    body = b""
    n = 0
    while n < w * h:
        k = int(rng.integers(1, 20))
        k = min(k, w * h - n)
        if rng.random() < 0.5:
            body += bytes([0x80 | (k - 1)]) + rng.integers(0, 256, size=4, dtype=np.uint8).tobytes()
        else:
            body += bytes([k - 1]) + rng.integers(0, 256, size=4 * k, dtype=np.uint8).tobytes()
        n += k
    files["rle32.tga"] = _tga(w, h, 10, 32, body)
    for name, data in files.items():
        p = str(tmp_path / name)
        with open(p, "wb") as f:
            f.write(data)
        _check(dec, stb, p)


def test_refused_and_unknown_files(dec, stb, tmp_path):
    from PIL import Image
    jpg = str(tmp_path / "t.jpg")
    Image.fromarray(np.zeros((8, 8, 3), np.uint8)).save(jpg)
    assert dec(jpg) == (None, "JPEG not supported")
    junk = str(tmp_path / "junk.bin")
    with open(junk, "wb") as f:
        f.write(b"\x07" * 40)
    assert dec(junk) == (None, "unknown image type")
    assert dec(str(tmp_path / "missing.png")) == (None, "can't fopen")
    if stb is not None:
        assert stb(junk)[1] == "unknown image type"
        assert stb(str(tmp_path / "missing.png"))[1] == "can't fopen"
