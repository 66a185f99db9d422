"""GPU draw pipeline vs the CPU oracle: colour, depth and stencil must be
bit-identical, and the covered/shaded fragment and primitive counts equal.

The oracle is itself pinned to the compiled reference renderer and to its
golden images (tests/test_oracle.py); the GPU results are additionally
checked against the golden PNGs with the reference harness' tolerance of 1
(tests/compare.py:8 of kitrofimov/srp)."""
import os

import ctypes

import numpy as np
import pytest

from scene_runner import REFERENCE_SCENES, REPO, write_noise_png

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(REPO, "tests", "golden", "references")


def _stats(mod, which):
    fn = {"oracle": "srpOracleGetStats", "gpu": "srpHipGetStats"}[which]
    return mod.stats(fn)


def _reset(mod, which):
    fn = {"oracle": "srpOracleResetStats", "gpu": "srpHipResetStats"}[which]
    mod.reset_stats(fn)


def depth_cull_on(gpu):
    return gpu.lib.srpHipGetTileDepthCull() != 0


def run_both(oracle, gpu, name, **kw):
    """Both frames and counters.  With tile-level depth culling on (the
    library default) the GPU's covered count omits fragments of triangles
    culled for a whole tile, so only shaded and primitive counts are
    compared then; the variants below turn culling off and compare the
    covered count too."""
    print(f"[parity] {name} {kw}", flush=True)
    _reset(oracle, "oracle")
    _reset(gpu, "gpu")
    a = oracle.run(name, **kw)
    b = gpu.run(name, **kw)
    sa, sb = list(_stats(oracle, "oracle")), list(_stats(gpu, "gpu"))
    if depth_cull_on(gpu):
        assert sb[0] <= sa[0], f"{name}: GPU covered {sb[0]} above the oracle's {sa[0]}"
        sa[0] = sb[0] = None
    return a, b, sa, sb


def assert_same(a, b, sa, sb, label):
    assert a.color.shape == b.color.shape, label
    dc = np.count_nonzero(a.color != b.color)
    dd = np.count_nonzero(a.depth.view(np.uint32) != b.depth.view(np.uint32))
    ds = np.count_nonzero(a.stencil != b.stencil)
    assert (dc, dd, ds) == (0, 0, 0), f"{label}: differing pixels colour={dc} depth={dd} stencil={ds}"
    assert sa[:3] == sb[:3], f"{label}: stats oracle={sa} gpu={sb} (covered, shaded, primitives)"


@pytest.mark.parametrize("name", REFERENCE_SCENES)
def test_reference_scene(oracle, gpu, name):
    a, b, sa, sb = run_both(oracle, gpu, name)
    assert_same(a, b, sa, sb, name)


@pytest.mark.parametrize("name", REFERENCE_SCENES)
def test_golden_image(gpu, name):
    from PIL import Image
    img = gpu.run(name).rgb().astype(np.int32)
    ref = np.asarray(Image.open(os.path.join(GOLDEN, name + ".png")).convert("RGB")).astype(np.int32)
    assert img.shape == ref.shape
    assert int(np.abs(img - ref).max()) <= 1


def test_cfg_tris_small(oracle, gpu):
    a, b, sa, sb = run_both(oracle, gpu, "cfg/tris", p0=300, seed=1)
    assert_same(a, b, sa, sb, "cfg/tris 300")


def test_cfg_tris_512(oracle, gpu):
    """Random clip-space triangles (2,000, 512x512, depth LESS): heavy clipping."""
    a, b, sa, sb = run_both(oracle, gpu, "cfg/tris", p0=2000, seed=0x5EED0001)
    assert_same(a, b, sa, sb, "cfg/tris 2000")


@pytest.fixture(scope="module")
def noise_tex(tmp_path_factory):
    return write_noise_png(str(tmp_path_factory.mktemp("tex") / "noise1024.png"), 1024, 0x5EED0002)


@pytest.mark.parametrize("shader", [0, 1, 2, 3, 4])
def test_cfg_terrain_small(oracle, gpu, shader):
    """Terrain heightfield 200x100 cells, 480x270: procedural material
    (shader 0) and the four G-buffer capture passes."""
    a, b, sa, sb = run_both(oracle, gpu, "cfg/terrain", width=480, height=270, p0=200, p1=100, p2=5, seed=shader)
    assert_same(a, b, sa, sb, f"cfg/terrain shader {shader}")


@pytest.mark.parametrize("pose", [0, 256])
def test_cfg_terrain_full_grid(oracle, gpu, pose):
    """The 1000x500-cell terrain (1,000,000 triangles) at 960x540."""
    a, b, sa, sb = run_both(oracle, gpu, "cfg/terrain", width=960, height=540, p2=pose)
    assert_same(a, b, sa, sb, f"cfg/terrain pose {pose}")


@pytest.mark.parametrize("batched,runs,box,cull", [(1, 0, 24, 1), (1, 1, 0, 1), (1, 0, 0, 1), (0, 1, 24, 1),
                                                   (1, 1, 8, 1), (1, 1, 24, 0), (1, 1, 0, 0)])
def test_raster_variants(oracle, gpu, batched, runs, box, cull):
    """Every single-draw raster variant gives the oracle's frames (and,
    with depth culling off, its covered counts)."""
    import ctypes
    lib = gpu.lib
    lib.srpHipSetRasterOptions.argtypes = [ctypes.c_int] * 3
    lib.srpHipSetRasterOptions(batched, runs, box)
    lib.srpHipSetTileDepthCull(cull)
    try:
        cases = [dict(name="cfg/tris", p0=2000, seed=0x5EED0001),
                 dict(name="cfg/mix", width=480, height=270, p0=20000, p1=4000, p2=4000, seed=3),
                 dict(name="cfg/terrain", width=480, height=270, p0=200, p1=100, p2=7),
                 dict(name="misc/stencil_test"), dict(name="interpolation/perspective")]
        cases += [dict(name="cfg/fuzz", seed=s) for s in range(0, 40, 3)]
        for c in cases:
            a, b, sa, sb = run_both(oracle, gpu, **c)
            assert_same(a, b, sa, sb, f"{c} variant {(batched, runs, box, cull)}")
    finally:
        lib.srpHipSetRasterOptions(1, 1, 24)
        lib.srpHipSetTileDepthCull(1)


@pytest.mark.parametrize("run", [0, 1, 7])
def test_tile_orders(oracle, gpu, noise_tex, run):
    """The frame does not depend on how a draw's tiles are spread over the
    XCDs (the default, 8, runs in every other test)."""
    lib = gpu.lib
    lib.srpHipSetTileOrder(run)
    try:
        assert lib.srpHipGetTileOrder() == run
        cases = [dict(name="cfg/mesh", width=1920, height=1080, p0=256, p1=512, tex_path=noise_tex),
                 dict(name="cfg/terrain", width=480, height=270, p0=200, p1=100, p2=7),
                 dict(name="cfg/tris", p0=2000, seed=0x5EED0001)]
        cases += [dict(name="cfg/fuzz", seed=s) for s in range(1, 40, 5)]
        for c in cases:
            a, b, sa, sb = run_both(oracle, gpu, **c)
            assert_same(a, b, sa, sb, f"{c} tile order {run}")
    finally:
        lib.srpHipSetTileOrder(8)


def test_cfg_mesh_small(oracle, gpu, noise_tex):
    a, b, sa, sb = run_both(oracle, gpu, "cfg/mesh", width=480, height=270, p0=32, p1=64, tex_path=noise_tex)
    assert_same(a, b, sa, sb, "cfg/mesh small")


def test_cfg_mesh_1080p(oracle, gpu, noise_tex):
    """The indexed textured UV sphere at 1920x1080 (256x512 segments)."""
    a, b, sa, sb = run_both(oracle, gpu, "cfg/mesh", width=1920, height=1080, p0=256, p1=512, tex_path=noise_tex)
    assert_same(a, b, sa, sb, "cfg/mesh 1080p")


def test_cfg_mix_small(oracle, gpu):
    a, b, sa, sb = run_both(oracle, gpu, "cfg/mix", width=480, height=270, p0=20000, p1=4000, p2=4000, seed=3)
    assert_same(a, b, sa, sb, "cfg/mix small")


@pytest.mark.parametrize("seed", range(0, 200))
def test_fuzz(oracle, gpu, seed):
    a, b, sa, sb = run_both(oracle, gpu, "cfg/fuzz", seed=seed)
    assert_same(a, b, sa, sb, f"fuzz seed {seed}")


def test_fuzz_sweep_raster_variants(oracle, gpu):
    """1,000 more randomised scenes, each under one of six rasteriser
    configurations in turn (batched with and without pixel runs, mask box
    24 / 8 / 0, tile depth culling on and off, and the simple per-pixel
    kernel)."""
    lib = gpu.lib
    lib.srpHipSetRasterOptions.argtypes = [ctypes.c_int] * 3
    variants = [(1, 1, 24, 1), (1, 0, 24, 1), (1, 1, 0, 1), (1, 1, 8, 1), (0, 1, 24, 1), (1, 1, 0, 0)]
    bad = []
    try:
        for seed in range(200, 1200):
            v = variants[seed % len(variants)]
            lib.srpHipSetRasterOptions(*v[:3])
            lib.srpHipSetTileDepthCull(v[3])
            a = oracle.run("cfg/fuzz", seed=seed)
            b = gpu.run("cfg/fuzz", seed=seed)
            if not (np.array_equal(a.color, b.color) and np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32))
                    and np.array_equal(a.stencil, b.stencil)):
                bad.append((seed, v))
    finally:
        lib.srpHipSetRasterOptions(1, 1, 24)
        lib.srpHipSetTileDepthCull(1)
    print(f"[parity] fuzz sweep: 1000 seeds, {len(bad)} differing", flush=True)
    assert not bad, f"differing (seed, options): {bad[:10]}"


# ---- against the compiled reference's fixtures (no oracle involved) ----

import ctypes  # noqa: E402
import hashlib  # noqa: E402
import json  # noqa: E402

from scene_runner import GPU_SO, ORACLE_SO, RES_DIR  # noqa: E402

HASHES = json.load(open(os.path.join(REPO, "tests", "golden", "reference_hashes.json")))


def _case_id(c):
    return ",".join(f"{k}={v}" for k, v in sorted(c.items()) if k != "tex_path")


def _digest(frame):
    return {
        "color": hashlib.sha256(frame.color.tobytes()).hexdigest(),
        "depth": hashlib.sha256(frame.depth.tobytes()).hexdigest(),
        "stencil": hashlib.sha256(frame.stencil.tobytes()).hexdigest(),
        "shape": list(frame.color.shape),
    }


@pytest.fixture(scope="module")
def fixture_tex(tmp_path_factory):
    return write_noise_png(str(tmp_path_factory.mktemp("ftex") / "noise1024.png"), 1024, HASHES["noise_seed"])


GPU_FIXTURE_CASES = [dict(name=n) for n in REFERENCE_SCENES] + [
    dict(name="cfg/tris", p0=300, seed=1),
    dict(name="cfg/tris", p0=2000, seed=0x5EED0001),
    dict(name="cfg/mesh", width=480, height=270, p0=32, p1=64),
    dict(name="cfg/mesh", width=1920, height=1080, p0=256, p1=512),
    dict(name="cfg/mix", width=480, height=270, p0=20000, p1=4000, p2=4000, seed=3),
] + [dict(name="cfg/fuzz", seed=s) for s in range(200)]


@pytest.mark.parametrize("case", GPU_FIXTURE_CASES, ids=_case_id)
def test_gpu_vs_reference_fixture(gpu, fixture_tex, case):
    c = dict(case)
    if c["name"] == "cfg/mesh":
        c["tex_path"] = fixture_tex
    print(f"[fixture] {_case_id(case)}", flush=True)
    assert _digest(gpu.run(**c)) == HASHES["cases"][_case_id(case)]


def _transcript(path):
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    buf = ctypes.create_string_buffer(1 << 16)
    lib.srpSceneMessages(RES_DIR.encode(), buf, 1 << 16)
    return buf.value.decode()


def test_gpu_messages_match_oracle(gpu):
    """Error/warning callbacks (OOB draws, excess vertices, unknown
    primitive, texture load failure, bad texture parameters)."""
    assert _transcript(GPU_SO) == _transcript(ORACLE_SO)


def test_deferred_frame_matches_oracle(fixture_tex):
    """Device-resident (deferred) framebuffer with device-side depth and
    stencil fills, as the benchmark runs it."""
    n = 480 * 270
    frames = []
    for path in (GPU_SO, ORACLE_SO):
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        lib.srpBenchMeshSetup.argtypes = [ctypes.c_int] * 4 + [ctypes.c_char_p]
        assert lib.srpBenchMeshSetup(480, 270, 64, 128, fixture_tex.encode()) == 0
        for _ in range(3):
            lib.srpBenchFrame()
        c, d, s = np.zeros(n, np.uint32), np.zeros(n, np.float32), np.zeros(n, np.uint8)
        lib.srpBenchRead(ctypes.c_void_p(c.ctypes.data), ctypes.c_void_p(d.ctypes.data), ctypes.c_void_p(s.ctypes.data))
        lib.srpBenchFinish()
        frames.append((c, d.view(np.uint32), s))
    for a, b in zip(*frames):
        assert np.array_equal(a, b)
