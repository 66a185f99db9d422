"""The device's float-only "roughly zero" test (include/srp/device/abi.h,
SRP_EPSILON_F) equals the reference's fabs(x) <= 1e-9 on a float promoted
to double (math/utils.h:45) for every float: the constant is the largest
float not above 1e-9, and the two tests agree on whole binades around it.
The rasteriser's one-compare edge test (SRP_EDGE_TL_THRESHOLD) equals the
reference's (l > 0) || (roughlyZero(l) && topLeft) (raster/triangle.c)."""
import os
import re

import numpy as np

from scene_runner import REPO


def _constant():
    text = open(os.path.join(REPO, "include", "srp", "device", "abi.h")).read()
    m = re.search(r"#define SRP_EPSILON_F (0x[0-9a-fA-Fp.+-]+)f", text)
    assert m, "SRP_EPSILON_F not found"
    return float.fromhex(m.group(1))


def test_constant_is_largest_float_not_above_epsilon():
    t = _constant()
    t32 = np.float32(t)
    assert float(t32) == t                      # exactly a float
    assert t <= 1e-9
    assert float(np.nextafter(t32, np.float32(1))) > 1e-9


def test_float_test_equals_double_test():
    t = np.float32(_constant())
    # every float of the binades [2^-31, 2^-29), both signs, plus specials
    bits = np.arange(0x30000000, 0x31000000, dtype=np.uint32)
    x = np.concatenate([bits.view(np.float32), -bits.view(np.float32),
                        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 1e-38], np.float32)])
    ref = np.abs(x.astype(np.float64)) <= 1e-9
    got = np.abs(x) <= t
    assert np.array_equal(ref, got)
    assert ref.sum() > 0 and (~ref).sum() > 0


def _threshold():
    text = open(os.path.join(REPO, "include", "srp", "device", "abi.h")).read()
    m = re.search(r"#define SRP_EDGE_TL_THRESHOLD \((-0x[0-9a-fA-Fp.+-]+)f\)", text)
    assert m, "SRP_EDGE_TL_THRESHOLD not found"
    return np.float32(float.fromhex(m.group(1)))


def test_edge_threshold_is_next_float_below_minus_epsilon():
    t = np.float32(_constant())
    thr = _threshold()
    assert thr < -t and np.nextafter(thr, np.float32(1)) == -t


def test_one_compare_edge_test_equals_reference_edge_test():
    """(l > 0) || (|l| <= eps && topLeft) equals l > (topLeft ? threshold : 0)
    for every float of the binades around eps (both signs), random bit
    patterns and the special values, for both edge kinds."""
    t = np.float32(_constant())
    thr = _threshold()
    bits = np.arange(0x2F000000, 0x31800000, dtype=np.uint32)
    rng = np.random.default_rng(5)
    x = np.concatenate([bits.view(np.float32), -bits.view(np.float32),
                        rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32),
                        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1e-38, -1e-38], np.float32)])
    with np.errstate(invalid="ignore"):
        for tl in (False, True):
            ref = (x > 0) | ((np.abs(x) <= t) & tl)
            got = x > (thr if tl else np.float32(0.0))
            assert np.array_equal(ref, got), tl
