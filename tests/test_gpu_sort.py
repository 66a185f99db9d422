"""The stable radix sort behind tile binning and per-pixel fragment lists.

Draws bin (tile, pair) items by sorting the tile keys with the pair index as
the value; the rasteriser then relies on each tile's pairs being in pair
(primitive) order, which is what a stable sort gives (the reference draws
primitives in ID order, pipeline/draw.c:84-165).  These tests sort host data
through the library's own sort scratch and compare with numpy's stable
argsort.  The sorts run back to back on that one scratch, so sums the sort
keeps between its passes must be back at zero for the next sort, and the
sizes cross every block layout: one block, the fused digit-offset passes up
to their 512-block limit, the scanned passes above it, and the long blocks
of more than 4 M items."""
import ctypes
import os

import numpy as np
import pytest

from scene_runner import REPO

LIB = os.path.join(REPO, "srp_amd", "libsrp_amd.so")

# (capacity, device count, key bits, key kind)
CASES = [
    (350_000, 350_000, 15, "tiles"),      # the 1080p mesh's pair count
    (1_000, 1_000, 15, "tiles"),          # one block
    (5_000, 3_001, 21, "uniform"),        # device count below the capacity, three passes
    (512 * 1024, 512 * 1024, 16, "few"),  # exactly the fused limit
    (440_000, 350_000, 15, "tiles"),      # capacity-sized grid, count from the device
    (600_000, 599_999, 12, "uniform"),    # scanned passes
    (5_000_000, 4_900_000, 15, "tiles"),  # long blocks
    (350_000, 350_000, 15, "tiles"),      # the first case again on the reused scratch
]


def _keys(rng, n, bits, kind):
    if kind == "uniform":
        return rng.integers(0, 1 << bits, size=n, dtype=np.uint32)
    if kind == "few":
        return rng.choice(np.array([3, 3, 200, 1 << 15, (1 << 16) - 1], np.uint32), size=n)
    # runs of neighbouring tiles, as consecutive triangles produce
    starts = rng.integers(0, (1 << bits) - 64, size=n // 8 + 1, dtype=np.uint32)
    return (np.repeat(starts, 8)[:n] + rng.integers(0, 3, size=n, dtype=np.uint32)) & ((1 << bits) - 1)


@pytest.mark.gpu
def test_radix_sort_matches_stable_argsort():
    lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_LOCAL)
    lib.srp_lib_radix_sort.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                       ctypes.c_int]
    rng = np.random.default_rng(20)
    for cap, count, bits, kind in CASES:
        print(f"[sort] cap={cap} count={count} bits={bits} keys={kind}", flush=True)
        keys = _keys(rng, cap, bits, kind)
        # bits above the sorted ones must ride along untouched
        keys |= rng.integers(0, 2, size=cap, dtype=np.uint32) << np.uint32(31)
        vals = np.arange(cap, dtype=np.uint32)
        k = keys.copy()
        v = vals.copy()
        lib.srp_lib_radix_sort(k.ctypes.data, v.ctypes.data, cap, count, bits)
        digit = keys[:count] & np.uint32((1 << (8 * ((bits + 7) // 8))) - 1)
        order = np.argsort(digit, kind="stable")
        assert np.array_equal(v[:count], vals[:count][order]), (cap, count, bits, kind)
        assert np.array_equal(k[:count], keys[:count][order]), (cap, count, bits, kind)
