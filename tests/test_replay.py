"""The rasteriser's barycentric replay (include/srp/device/replay.h) equals
the reference's sequential float additions bit for bit: a host build of the
same header against the naive loop on 3 million random and adversarial
cases."""
import os
import shutil
import subprocess

import pytest

from scene_runner import REPO


def test_fast_replay_equals_sequential_additions(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "replay_check")
    subprocess.run([cc, "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "native", "replay_check.c"), "-lm", "-o", exe], check=True)
    r = subprocess.run([exe, "3000000"], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().splitlines()[-1].startswith("0 of ")
