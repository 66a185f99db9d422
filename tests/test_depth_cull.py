"""Tile-level depth culling (include/srp/device/depth_cull.h) skips a
triangle for a whole tile only when none of its fragments there can pass
the depth test.  That rests on srpTileDepthRange bounding the depth the
reference computes at every pixel (sequential barycentric walk, three-term
depth sum).  A host build checks the range against every pixel of 200,000
random, adversarial (long walks, near-cancelling depth planes, zero steps)
and realistically set-up triangles; the same check with the error term
removed, or with the walk's error bound shrunk, fails (kept here as the
mutants that show the check bites).  The GPU frames with culling on equal
the oracle's in tests/test_gpu_parity.py (fuzz sweep, variants, full-size
primitive mix)."""
import os
import re
import shutil
import subprocess

import pytest

from scene_runner import REPO

SRC = os.path.join(REPO, "tests", "native", "depth_cull_check.c")
HEADER = os.path.join(REPO, "include", "srp", "device", "depth_cull.h")


def _build(tmp_path, include_dirs, name):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    exe = str(tmp_path / name)
    cmd = [cc, "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math"]
    for d in include_dirs:
        cmd += ["-I", d]
    subprocess.run(cmd + [SRC, "-lm", "-o", exe], check=True)
    return exe


def _run(exe, n):
    r = subprocess.run([exe, str(n)], stdout=subprocess.PIPE, text=True)
    last = r.stdout.strip().splitlines()[-1]
    return r.returncode, int(last.split()[0]), last


def test_depth_range_holds_every_pixel(tmp_path):
    exe = _build(tmp_path, [os.path.join(REPO, "include")], "dcc")
    rc, bad, line = _run(exe, 200000)
    print(line)
    assert rc == 0 and bad == 0


@pytest.mark.parametrize("mutation", [
    (r"err \+= fabs\(a\) \* \(em \+ g3 \* \(m \+ em\) \+ 1e-12 \* m\);", "err += 0.0;"),
    (r"const double gN = n \* u / \(1\.0 - n \* u\);", "const double gN = 3.0 * u;"),
])
def test_check_catches_a_weakened_bound(tmp_path, mutation):
    text = open(HEADER).read()
    mutated, k = re.subn(mutation[0], mutation[1], text)
    assert k == 1
    d = tmp_path / "inc" / "srp" / "device"
    d.mkdir(parents=True)
    (d / "depth_cull.h").write_text(mutated)
    exe = _build(tmp_path, [str(tmp_path / "inc"), os.path.join(REPO, "include")], "dcc_mut")
    rc, bad, line = _run(exe, 200000)
    print(line)
    assert rc == 1 and bad > 0
