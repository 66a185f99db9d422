"""CPU checks of bench.py's command-line contract and derived figures (no
GPU): a request for more than one GPU is refused with exit status 2 before
any device work; the secondary VALU figure is linear in its per-frame
counts and expressed against the FP32 vector peak; the host description
names the cores visible."""
import os
import subprocess
import sys

import pytest

from scene_runner import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", n],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_valu_figure_is_linear_in_counts():
    a = bench.valu_figure(1000, 200, 100, 0.5, 1.0)
    b = bench.valu_figure(3000, 50, 70, 0.5, 1.0)
    ab = bench.valu_figure(4000, 250, 170, 0.5, 1.0)
    assert ab["ops_per_frame"] == a["ops_per_frame"] + b["ops_per_frame"]
    # one bounding-box pixel costs the three barycentric steps and nothing else
    only_box = bench.valu_figure(1_000_000, 0, 0, 1.0, 2.0)
    assert only_box["ops_per_frame"] == 3_000_000
    assert only_box["achieved"] == pytest.approx(3e6 / 1e-3 / 1e12)
    assert only_box["frame_achieved"] == pytest.approx(3e6 / 2e-3 / 1e12)
    assert only_box["peak"] == 157.3 and only_box["unit"] == "TFLOP/s"
    assert only_box["frac"] == pytest.approx(only_box["achieved"] / 157.3, abs=5e-7)   # rounded to 6 places
    assert bench.valu_figure(10, 1, 1, 0.0, 1.0)["achieved"] == 0.0


def test_host_cpu_reports_cores():
    model, n = bench.host_cpu()
    assert n == os.cpu_count() and n >= 1
    assert model is None or isinstance(model, str)
