"""A single draw's emitting half is sized from the same draw's previous
counts and enqueued before the host has read the new ones; the library
checks them on the device (the half is skipped when they do not fit) and
on the host at its next call, which repeats the half with buffers for the
counts.  These tests issue one draw 15 times with counts that grow and
shrink between calls (scene spec/grow: filled, line and point polygon
modes), on a synchronous and on a device-resident (deferred) framebuffer,
and compare every result with the oracle, which draws as the reference does
(tests/test_oracle.py pins it to the compiled reference on this scene)."""
import pytest

from test_gpu_parity import assert_same, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deferred", [0, 1])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_repeated_draw_with_changing_counts(oracle, gpu, seed, deferred):
    a, b, sa, sb = run_both(oracle, gpu, "spec/grow", p0=300, p1=deferred, seed=seed)
    assert_same(a, b, sa, sb, f"spec/grow seed={seed} deferred={deferred}")


def test_repeated_frames_of_one_scene_stay_equal(oracle, gpu):
    """The same frame drawn three times in a row (the second and third
    draws take the speculative path) is the oracle's every time."""
    first = oracle.run("spec/grow", p0=200, seed=9)
    for k in range(3):
        print(f"[speculative] repeat {k}", flush=True)
        g = gpu.run("spec/grow", p0=200, p1=1, seed=9)
        assert_same(first, g, [0, 0, 0], [0, 0, 0], f"repeat {k}")
