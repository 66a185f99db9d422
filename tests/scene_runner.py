"""ctypes driver for the scene modules (oracle, compiled reference, GPU).

Each module exports srpSceneRun(name, SceneArgs*, SceneOut*) (tests/scenes/
scenes.h); this wrapper allocates the output buffers and returns numpy
arrays of colour (uint32 R<<24|G<<16|B<<8|A), depth (float32) and stencil
(uint8), row-major H x W.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RES_DIR = os.path.join(REPO, "tests", "golden", "res")
ORACLE_SO = os.path.join(REPO, "oracle", "build", "libscenes_oracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libscenes_ref.so")
GPU_SO = os.path.join(REPO, "tests", "scenes", "build", "libscenes_amd.so")

REFERENCE_SCENES = [
    "primitives/points", "primitives/lines", "primitives/line_strip", "primitives/line_loop",
    "primitives/triangle_strip", "primitives/triangle_fan",
    "polygon_modes/line", "polygon_modes/point",
    "clipping/triangle", "clipping/line", "clipping/point",
    "interpolation/perspective", "interpolation/affine", "interpolation/flat",
    "misc/depth_test", "misc/stencil_test", "misc/scissor_test", "misc/rasterization_rules",
]


class SceneArgs(ctypes.Structure):
    _fields_ = [
        ("resDir", ctypes.c_char_p), ("texPath", ctypes.c_char_p),
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("p0", ctypes.c_int), ("p1", ctypes.c_int), ("p2", ctypes.c_int),
        ("f0", ctypes.c_float), ("seed", ctypes.c_uint64),
    ]


class SceneOut(ctypes.Structure):
    _fields_ = [
        ("color", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("stencil", ctypes.c_void_p),
        ("capacity", ctypes.c_size_t), ("width", ctypes.c_size_t), ("height", ctypes.c_size_t),
    ]


class Frame:
    def __init__(self, color, depth, stencil):
        self.color, self.depth, self.stencil = color, depth, stencil

    def rgb(self):
        """H x W x 3 uint8 (the reference harness writes RGB with alpha 255)."""
        c = self.color
        return np.stack([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255], axis=-1).astype(np.uint8)


class SceneModule:
    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.lib.srpSceneRun.argtypes = [ctypes.c_char_p, ctypes.POINTER(SceneArgs), ctypes.POINTER(SceneOut)]
        self.lib.srpSceneRun.restype = ctypes.c_int

    def run(self, name, width=0, height=0, p0=0, p1=0, p2=0, f0=0.0, seed=0, tex_path=None,
            capacity=None):
        cap = capacity or max(512 * 512, (width or 1920) * (height or 1080), 256 * 256)
        color = np.zeros(cap, np.uint32)
        depth = np.zeros(cap, np.float32)
        stencil = np.zeros(cap, np.uint8)
        args = SceneArgs(RES_DIR.encode(), tex_path.encode() if tex_path else None,
                         width, height, p0, p1, p2, f0, seed)
        out = SceneOut(color.ctypes.data, depth.ctypes.data, stencil.ctypes.data, cap, 0, 0)
        rc = self.lib.srpSceneRun(name.encode(), ctypes.byref(args), ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(f"scene {name} failed with code {rc}")
        n = out.width * out.height
        shape = (out.height, out.width)
        return Frame(color[:n].reshape(shape).copy(), depth[:n].reshape(shape).copy(),
                     stencil[:n].reshape(shape).copy())

    def stats(self, fn="srpOracleGetStats"):
        arr = (ctypes.c_uint64 * 4)()
        getattr(self.lib, fn)(arr)
        return list(arr)

    def reset_stats(self, fn="srpOracleResetStats"):
        getattr(self.lib, fn)()


def write_noise_png(path, size, seed):
    """Seeded uniform-noise RGB texture (the indexed-mesh workload's)."""
    from PIL import Image
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, size=(size, size, 3), dtype=np.uint8)
    Image.fromarray(img, "RGB").save(path)
    return path
