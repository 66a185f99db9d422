"""The image harness (tests/image_harness.py): frames written by the
library's PNG writer read back through PIL exactly (RGB as stored, alpha
255); a golden comparison passes within the tolerance of 1 and fails beyond
it, writing the x20 difference and failing-pixel mask images; goldens
regenerated from the compiled reference equal the reference's own.  The
GPU variant round-trips a GPU frame (srpFramebufferSaveImage)."""
import ctypes
import os

import numpy as np
import pytest
from PIL import Image

import image_harness as ih
from scene_runner import GPU_SO, REF_SO, REFERENCE_SCENES, SceneModule


def _rgba(color):
    c = color.astype(np.uint32)
    return np.stack([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255, np.full_like(c, 255)], -1).astype(np.uint8)


@pytest.mark.parametrize("scene", ["misc/stencil_test", "interpolation/perspective", "primitives/lines"])
def test_oracle_frame_round_trips_and_matches_golden(oracle, tmp_path, scene):
    frame = oracle.run(scene)
    path = ih.save_frame(frame.color, str(tmp_path / (scene.replace("/", "_") + ".png")))
    got = np.asarray(Image.open(path))
    assert got.shape == frame.color.shape + (4,) and np.array_equal(got, _rgba(frame.color))
    ok, rep = ih.compare(path, os.path.join(ih.GOLDEN_DIR, scene + ".png"))
    assert ok, rep


def test_compare_fails_beyond_tolerance_with_diff_and_mask(oracle, tmp_path):
    frame = oracle.run("misc/depth_test")
    ref = ih.save_frame(frame.color, str(tmp_path / "ref.png"))
    c = frame.color.copy()
    c[10, 20] += 1 << 8            # blue +1: within the tolerance
    ok, rep = ih.compare(ih.save_frame(c, str(tmp_path / "near.png")), ref)
    assert ok and rep["max_diff"] == 1
    c[30, 40] += 5 << 24           # red +5
    c[31, 41] ^= 0xFF << 16        # green flipped
    out = ih.save_frame(c, str(tmp_path / "far.png"))
    ok, rep = ih.compare(out, ref, out)
    assert not ok and rep["differing_pixels"] == 2 and rep["first_failing_pixel"] == (40, 30)
    diff = np.asarray(Image.open(rep["diff_image"]))
    mask = np.asarray(Image.open(rep["mask_image"]))
    assert diff[30, 40, 0] == 100 and diff[10, 20, 2] == 20 and diff[0, 0].max() == 0
    assert mask[30, 40] == 255 and mask[31, 41] == 255 and mask[10, 20] == 0 and int(mask.sum()) == 2 * 255


def test_regenerated_goldens_equal_the_references(tmp_path):
    if not os.path.exists(REF_SO):
        pytest.skip("compiled reference not built (oracle/_ref)")
    ref = SceneModule(REF_SO)
    scenes = [s for s in REFERENCE_SCENES if s != "primitives/lines"]   # its committed golden is stale
    for path, scene in zip(ih.regen(scenes, module=ref, dest=str(tmp_path)), scenes):
        ok, rep = ih.compare(path, os.path.join(ih.GOLDEN_DIR, scene + ".png"))
        assert ok and rep["max_diff"] == 0, (scene, rep)


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["misc/stencil_test", "clipping/triangle", "interpolation/flat"])
def test_gpu_frame_round_trips_and_matches_golden(gpu, tmp_path, scene):
    print(f"[image] {scene}", flush=True)
    ok, rep = ih.check(scene, out_dir=str(tmp_path), module=gpu)
    assert ok, rep
    frame = gpu.run(scene)
    got = np.asarray(Image.open(str(tmp_path / (scene + ".png"))))
    assert np.array_equal(got, _rgba(frame.color))


@pytest.mark.gpu
def test_framebuffer_save_image_of_a_deferred_frame(tmp_path):
    """srpFramebufferSaveImage on a device-resident framebuffer writes the
    device copy (it downloads first)."""
    lib = ctypes.CDLL(GPU_SO, mode=ctypes.RTLD_LOCAL)
    lib.srpBenchTrisSetup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    lib.srpBenchFramebuffer.restype = ctypes.c_void_p
    lib.srpFramebufferSaveImage.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.srpBenchRead.argtypes = [ctypes.c_void_p] * 3
    assert lib.srpBenchTrisSetup(256, 192, 300, 7) == 0
    try:
        lib.srpBenchFrame()
        path = str(tmp_path / "tris.png")
        assert lib.srpFramebufferSaveImage(lib.srpBenchFramebuffer(), path.encode()) == 0
        c = np.zeros(256 * 192, np.uint32)
        d = np.zeros(c.size, np.float32)
        s = np.zeros(c.size, np.uint8)
        lib.srpBenchRead(c.ctypes.data, d.ctypes.data, s.ctypes.data)
        assert np.array_equal(np.asarray(Image.open(path)), _rgba(c.reshape(192, 256)))
        assert len(np.unique(c)) > 100
    finally:
        lib.srpBenchFinish()
