"""Image-based checks in the style of the reference's test harness
(kitrofimov/srp tests/utils/save.c:4-33, tests/compare.py:8-80,
tests/gen_ref.py:7-26): a scene's frame is written as an RGBA PNG with alpha
255, compared with its golden image channel by channel with a tolerance of
1; on failure an amplified difference image (x20, `<out>.diff.png`) and a
mask of the failing pixels (`<out>.mask.png`) are written next to it.
Goldens are regenerated from the reference renderer compiled in place
(oracle/_ref) rather than from the build under test.

    python tests/image_harness.py check misc/stencil_test [--renderer gpu|oracle|ref] [--out DIR]
    python tests/image_harness.py regen all [--renderer ref] [--dest DIR]
"""
import argparse
import ctypes
import os
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from scene_runner import GPU_SO, ORACLE_SO, REF_SO, REFERENCE_SCENES, REPO, SceneModule  # noqa: E402

GOLDEN_DIR = os.path.join(REPO, "tests", "golden", "references")
LIB = os.path.join(REPO, "srp_amd", "libsrp_amd.so")
TOLERANCE = 1
DIFF_SCALE = 20
RENDERERS = {"gpu": GPU_SO, "oracle": ORACLE_SO, "ref": REF_SO}


def _writer():
    lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_LOCAL)
    lib.srpPngSaveColorBuffer.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.srpPngSaveColorBuffer.restype = ctypes.c_int
    return lib


def save_frame(color, path):
    """The colour buffer (H x W u32, R in the top byte) as an RGBA PNG with
    alpha 255, by the library's PNG writer."""
    c = np.ascontiguousarray(color, dtype=np.uint32)
    h, w = c.shape
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    if _writer().srpPngSaveColorBuffer(path.encode(), w, h, c.ctypes.data) != 0:
        raise OSError(f"could not write {path}")
    return path


def _load(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGBA")).astype(np.float64)


def compare(path, ref_path, out_path=None):
    """Channel-wise comparison with TOLERANCE.  Returns (ok, report dict);
    on failure writes <out>.diff.png and <out>.mask.png."""
    from PIL import Image
    a, b = _load(path), _load(ref_path)
    if a.shape != b.shape:
        return False, {"error": f"size mismatch {a.shape} vs {b.shape}"}
    diff = np.abs(a - b)
    bad = np.any(diff > TOLERANCE, axis=2)
    rep = {"max_diff": float(diff.max()), "mse": float(np.mean(diff ** 2)), "differing_pixels": int(bad.sum()),
           "pixels": int(bad.size)}
    if rep["max_diff"] <= TOLERANCE:
        return True, rep
    y, x = np.argwhere(bad)[0]
    rep["first_failing_pixel"] = (int(x), int(y))
    base = os.path.splitext(out_path or path)[0]
    rep["diff_image"] = base + ".diff.png"
    rep["mask_image"] = base + ".mask.png"
    Image.fromarray(np.clip(diff * DIFF_SCALE, 0, 255).astype(np.uint8)).save(rep["diff_image"])
    Image.fromarray(bad.astype(np.uint8) * 255).save(rep["mask_image"])
    return False, rep


def check(scene, renderer="gpu", out_dir=None, module=None):
    """Render `scene`, write its frame and compare it with the golden."""
    mod = module or SceneModule(RENDERERS[renderer])
    frame = mod.run(scene)
    out_dir = out_dir or os.environ.get("SRP_IMAGE_DIR") or os.path.join(tempfile.gettempdir(), "srp_images")
    out = save_frame(frame.color, os.path.join(out_dir, scene + ".png"))
    return compare(out, os.path.join(GOLDEN_DIR, scene + ".png"), out)


def regen(scenes, renderer="ref", dest=None, module=None):
    """Goldens from the compiled reference (by default): one PNG per scene
    under dest (the reference's tests/references layout)."""
    mod = module or SceneModule(RENDERERS[renderer])
    dest = dest or GOLDEN_DIR
    written = []
    for s in scenes:
        written.append(save_frame(mod.run(s).color, os.path.join(dest, s + ".png")))
    return written


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("check")
    c.add_argument("scene")
    c.add_argument("--renderer", choices=sorted(RENDERERS), default="gpu")
    c.add_argument("--out", default=None)
    r = sub.add_parser("regen")
    r.add_argument("scene", help="a scene name or 'all'")
    r.add_argument("--renderer", choices=sorted(RENDERERS), default="ref")
    r.add_argument("--dest", default=None)
    a = ap.parse_args()
    if a.cmd == "check":
        ok, rep = check(a.scene, a.renderer, a.out)
        print(("OK " if ok else "FAIL ") + " ".join(f"{k}={v}" for k, v in rep.items()))
        return 0 if ok else 1
    scenes = REFERENCE_SCENES if a.scene == "all" else [a.scene]
    for p in regen(scenes, a.renderer, a.dest):
        print("wrote", p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
