"""Full-size workloads whose CPU frames take minutes: the GPU frames must
hash to what the reference renderer (compiled in place, oracle/_ref)
produced for the same scene programs (tests/golden/make_fixtures.py
--heavy).  The oracle equals the reference on these too (recorded when the
fixtures were made; the same scene programs run in tests/test_oracle.py at
smaller sizes)."""
import json
import os

import pytest

from scene_runner import REPO

pytestmark = pytest.mark.gpu

HEAVY = os.path.join(REPO, "tests", "golden", "reference_hashes_heavy.json")


def _cases():
    if not os.path.exists(HEAVY):
        return []
    with open(HEAVY) as f:
        return sorted(json.load(f)["cases"].items())


def _args(case_id):
    kw = {}
    for part in case_id.split(","):
        k, v = part.split("=", 1)
        kw[k] = v if k == "name" else int(v)
    return kw


@pytest.mark.parametrize("case_id,want", _cases(), ids=[c for c, _ in _cases()])
def test_full_size_frame_matches_reference(gpu, case_id, want):
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    from make_fixtures import digest
    print(f"[heavy] {case_id}", flush=True)
    got = digest(gpu.run(**_args(case_id)))
    assert got == want
