"""Per-draw limits and the exact totals they are checked on.

A draw whose counting pass totals (tile entries or fragments, primitives,
vertices) do not fit the emitting pass's 32-bit offsets is refused through
the message callback before anything is sized or emitted (the reference's
convention for a draw it cannot perform, pipeline/draw.c:167-189).  The CPU
test drives the host-side decision with synthetic totals beyond 2^32; the
GPU test checks that the scan delivers those totals exactly (64-bit) where
its 32-bit offsets wrap.  No over-limit draw is run on the GPU."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from scene_runner import REPO

LIB = os.path.join(REPO, "srp_amd", "libsrp_amd.so")


def test_draw_limit_decisions_beyond_2_32(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path / "sizing_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "srp_amd", "csrc", "lib"),
                    os.path.join(REPO, "tests", "native", "sizing_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok")


@pytest.mark.gpu
@pytest.mark.parametrize("n,arrays", [(1 << 20, 3), (17 * (1 << 20), 1), (1, 3), (2049, 2)])
def test_scan_totals_are_exact_beyond_2_32(n, arrays):
    """Counts whose sums pass 2^32 (up to ~2^44): offsets and 32-bit totals
    wrap as unsigned arithmetic does, the 64-bit totals are exact (both the
    two-launch path and, for 17 M entries, the recursive one)."""
    print(f"[limits] scan n={n} arrays={arrays}", flush=True)
    lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_LOCAL)
    lib.srp_lib_scan_exact.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(n + arrays)
    counts = rng.integers(0, 1 << 24, size=(arrays, n), dtype=np.uint64)
    counts[:, ::97] = (1 << 32) - 1 - rng.integers(0, 1000, size=counts[:, ::97].shape, dtype=np.uint64)
    c32 = counts.astype(np.uint32)
    offs = np.zeros_like(c32)
    t32 = np.zeros(arrays, np.uint32)
    t64 = np.zeros(arrays, np.uint64)
    lib.srp_lib_scan_exact(c32.ctypes.data, n, arrays, offs.ctypes.data, t32.ctypes.data, t64.ctypes.data)
    exact = counts.sum(axis=1, dtype=np.uint64)
    assert np.array_equal(t64, exact)
    assert np.array_equal(t32, (exact & 0xFFFFFFFF).astype(np.uint32))
    want = np.concatenate([np.zeros((arrays, 1), np.uint64), np.cumsum(counts, axis=1, dtype=np.uint64)[:, :-1]],
                          axis=1)
    assert np.array_equal(offs, (want & 0xFFFFFFFF).astype(np.uint32))
    assert int(exact.max()) > (1 << 32) or n < 100
