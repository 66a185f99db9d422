"""Writes tests/golden/reference_hashes.json from the reference renderer
compiled in place (oracle/_ref/libscenes_ref.so, built by `make -C oracle
ref` from /root/reference).  For each case: SHA-256 of the colour, depth and
stencil buffers the reference produces.  These pin the oracle (CPU tests)
and the GPU path (GPU tests) without the reference itself being present.

    python tests/golden/make_fixtures.py            # reference_hashes.json
    python tests/golden/make_fixtures.py --heavy    # reference_hashes_heavy.json

The heavy cases are the full-size workloads whose CPU frames take minutes
(the primitive mix at 1920x1080 covers 3.9e9 fragments); the GPU tests
compare their frames with these hashes directly.  `--heavy` computes only
the cases the file does not hold yet (`--redo` recomputes all of them).
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from scene_runner import REF_SO, REFERENCE_SCENES, SceneModule, write_noise_png  # noqa: E402

NOISE_SEED = 0x5EED0002


def cases(tex):
    out = [dict(name=n) for n in REFERENCE_SCENES]
    out += [
        dict(name="cfg/tris", p0=300, seed=1),
        dict(name="cfg/tris", p0=2000, seed=0x5EED0001),
        dict(name="cfg/mesh", width=480, height=270, p0=32, p1=64, tex_path=tex),
        dict(name="cfg/mesh", width=1920, height=1080, p0=256, p1=512, tex_path=tex),
        dict(name="cfg/mix", width=480, height=270, p0=20000, p1=4000, p2=4000, seed=3),
    ]
    out += [dict(name="cfg/fuzz", seed=s) for s in range(200)]
    return out


def heavy_cases():
    return [
        dict(name="cfg/mix", width=1920, height=1080, p0=500000, p1=100000, p2=100000, seed=0x5EED0003),
        dict(name="cfg/terrain", width=1920, height=1080, p2=0),
        dict(name="cfg/terrain", width=1920, height=1080, p2=300),
    ] + [
        # the 1080p terrain from more of its camera path: training poses,
        # the timed pose (256), and test and validation poses of the
        # learned shader
        dict(name="cfg/terrain", width=1920, height=1080, p2=p)
        for p in (37, 64, 128, 192, 256, 280, 319, 333, 351)
    ]


def case_id(c):
    return ",".join(f"{k}={v}" for k, v in sorted(c.items()) if k != "tex_path")


def digest(frame):
    return {
        "color": hashlib.sha256(frame.color.tobytes()).hexdigest(),
        "depth": hashlib.sha256(frame.depth.tobytes()).hexdigest(),
        "stencil": hashlib.sha256(frame.stencil.tobytes()).hexdigest(),
        "shape": list(frame.color.shape),
    }


def main():
    ref = SceneModule(REF_SO)
    heavy = "--heavy" in sys.argv
    table = {}
    if heavy:
        out = "reference_hashes_heavy.json"
        path = os.path.join(HERE, out)
        if "--redo" not in sys.argv and os.path.exists(path):
            with open(path) as f:
                table = json.load(f)["cases"]
        for c in heavy_cases():
            if case_id(c) in table:
                continue
            ref.reset_stats("srpRefResetStats") if hasattr(ref.lib, "srpRefResetStats") else None
            table[case_id(c)] = digest(ref.run(**c))
            print("done", case_id(c), flush=True)
        meta = {}
    else:
        tex = write_noise_png("/tmp/srp_fixture_noise1024.png", 1024, NOISE_SEED)
        for c in cases(tex):
            table[case_id(c)] = digest(ref.run(**c))
        out = "reference_hashes.json"
        meta = {"noise_seed": NOISE_SEED}
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(dict(meta, cases=table), f, indent=1, sort_keys=True)
    print(f"wrote {len(table)} cases to {out}")


if __name__ == "__main__":
    main()
