"""The allocation-free PSNR evaluator (srpNeuralCompareNoAlloc, through the
C ABI) against srpNeuralCompare on the same framebuffers: identical PSNR,
largest difference and pixel count, on two calls in a row, the second on a
smaller frame (the static scratch keeps the first call's partials; the
second must not see them), and once more on the first size."""
import ctypes

import numpy as np
import pytest

from neural_runner import LIB_SO

pytestmark = pytest.mark.gpu


class _Framebuffer(ctypes.Structure):
    """SRPFramebuffer of include/srp/framebuffer.h."""
    _fields_ = [("width", ctypes.c_size_t), ("height", ctypes.c_size_t), ("size", ctypes.c_size_t),
                ("color", ctypes.POINTER(ctypes.c_uint32)), ("depth", ctypes.POINTER(ctypes.c_float)),
                ("stencil", ctypes.POINTER(ctypes.c_uint8))]


def _lib():
    L = ctypes.CDLL(LIB_SO, mode=ctypes.RTLD_LOCAL)
    L.srpNewFramebuffer.restype = ctypes.POINTER(_Framebuffer)
    L.srpNewFramebuffer.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    L.srpFreeFramebuffer.argtypes = [ctypes.POINTER(_Framebuffer)]
    for name in ("srpNeuralCompare", "srpNeuralCompareNoAlloc"):
        getattr(L, name).argtypes = [ctypes.POINTER(_Framebuffer)] * 3 + [ctypes.c_void_p] * 3
        getattr(L, name).restype = None
    return L


def _frames(L, w, h, covered, seed):
    rng = np.random.default_rng(seed)
    fbs = [L.srpNewFramebuffer(w, h) for _ in range(3)]
    n = w * h
    for fb in fbs:
        np.ctypeslib.as_array(fb.contents.color, (n,))[:] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        np.ctypeslib.as_array(fb.contents.depth, (n,))[:] = np.where(
            rng.uniform(size=n) < covered, rng.uniform(0, 0.999, n), 1.0).astype(np.float32)
    ca = np.ctypeslib.as_array(fbs[0].contents.color, (n,))
    cb = np.ctypeslib.as_array(fbs[1].contents.color, (n,))
    cb[:] = (ca.astype(np.int64) ^ (rng.integers(0, 8, n) << (8 * rng.integers(1, 4, n)))).astype(np.uint32)
    return fbs


def _compare(L, name, fbs):
    p, m, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
    getattr(L, name)(fbs[0], fbs[1], fbs[2], ctypes.byref(p), ctypes.byref(m), ctypes.byref(n))
    return p.value, m.value, n.value


def test_no_alloc_compare_equals_compare():
    """Two no-alloc calls in a row (1080p, then a smaller ragged frame), then
    the existing entry point on both; then a third no-alloc call on a tiny
    frame after a large one."""
    L = _lib()
    sets = [_frames(L, w, h, cov, 1000 + k) for k, (w, h, cov) in
            enumerate([(1920, 1080, 0.7), (333, 77, 0.4), (1920, 1080, 0.2), (5, 3, 1.0)])]
    try:
        print("[no-alloc] 1920x1080 then 333x77", flush=True)
        got = [_compare(L, "srpNeuralCompareNoAlloc", sets[0]), _compare(L, "srpNeuralCompareNoAlloc", sets[1])]
        want = [_compare(L, "srpNeuralCompare", sets[0]), _compare(L, "srpNeuralCompare", sets[1])]
        assert got == want
        print("[no-alloc] 1920x1080 then 5x3", flush=True)
        got = [_compare(L, "srpNeuralCompareNoAlloc", sets[2]), _compare(L, "srpNeuralCompareNoAlloc", sets[3])]
        want = [_compare(L, "srpNeuralCompare", sets[2]), _compare(L, "srpNeuralCompare", sets[3])]
        assert got == want
        assert all(r[2] > 0 for r in got)
    finally:
        for fbs in sets:
            for fb in fbs:
                L.srpFreeFramebuffer(fb)
