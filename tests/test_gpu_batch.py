"""Batched rendering of independent scenes (srpNewBatch / srpBatchRender):
one persistent rasteriser launch for every scene's tiles, HIP graph
capture and replay, and the autotuner.  Every path must give frames
bit-identical to drawing the scenes one at a time, on the GPU and in the
CPU oracle."""
import pytest

from batch_runner import C4_SEED, BatchScenes, frames_equal
from scene_runner import GPU_SO, ORACLE_SO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_batch():
    return BatchScenes(GPU_SO, gpu=True)


@pytest.fixture(scope="module")
def oracle_batch():
    return BatchScenes(ORACLE_SO, gpu=False)


def oracle_frames(ob, n, scale, seed):
    ob.setup(n, scale, seed, batched=False)
    try:
        ob.frame(False)
        return ob.read_all()
    finally:
        ob.finish()


@pytest.mark.parametrize("n,scale,seed", [(12, 0.05, C4_SEED), (24, 0.02, 7), (64, 0.01, 11)])
def test_batch_matches_draw_by_draw_and_oracle(gpu_batch, oracle_batch, n, scale, seed):
    print(f"[batch] {n} scenes scale {scale} seed {seed}", flush=True)
    ref = oracle_frames(oracle_batch, n, scale, seed)
    gb = gpu_batch
    gb.setup(n, scale, seed, batched=True)
    try:
        gb.frame(True)
        batched = gb.read_all()
        gb.frame(False)
        single = gb.read_all()
    finally:
        gb.finish()
    assert frames_equal(single, ref) == []
    assert frames_equal(batched, ref) == []


@pytest.mark.parametrize("tile", [8, 16])
def test_full_scale_batch_autotuned_matches_oracle(gpu_batch, oracle_batch, tile):
    """The batched-scenes workload at full size (64 scenes, 1.86 M
    triangles, framebuffers up to 2048 x 2048): autotuned, then at the best
    waves/CU of each tile size, captured in a graph and replayed, every
    scene's frame equals the oracle's (about 4 s of CPU)."""
    print(f"[batch] full scale, tile {tile}", flush=True)
    ref = oracle_frames(oracle_batch, 64, 1.0, C4_SEED)
    gb = gpu_batch
    gb.setup(64, 1.0, C4_SEED, batched=True)
    try:
        ms, best = gb.autotune(1)
        print(f"[batch] autotune {ms:.2f} ms: {best}", flush=True)
        waves, box, xcd, runs, _ = best
        gb.set_tuning(waves, box, xcd, runs, tile)
        gb.frame(True)
        assert frames_equal(gb.read_all(), ref) == []
        assert gb.capture() == 0
        gb.frame(True)
        gb.frame(True)
        assert frames_equal(gb.read_all(), ref) == []
    finally:
        gb.finish()


def test_graph_replays_match_eager(gpu_batch):
    """The captured frame, replayed several times with eager work (draw by
    draw frames) in between, equals the eager batch frame every time."""
    gb = gpu_batch
    gb.setup(16, 0.05, 3, batched=True)
    try:
        gb.frame(True)
        eager = gb.read_all()
        assert gb.capture() == 0
        for r in range(3):
            gb.frame(True)          # graph replay
            got = gb.read_all()
            assert frames_equal(got, eager) == [], f"replay {r}"
            gb.frame(False)         # eager single draws in between
    finally:
        gb.finish()


def test_tunings_give_identical_frames(gpu_batch):
    gb = gpu_batch
    gb.setup(16, 0.05, 5, batched=True)
    try:
        gb.frame(True)
        base = gb.read_all()
        for waves, box, xcd, runs, tile in [(1, 0, 0, 0, 8), (2, 8, 1, 1, 8), (16, 16, 0, 0, 8), (32, 24, 1, 1, 8),
                                            (4, 0, 1, 1, 16), (8, 24, 0, 0, 8), (16, 24, 1, 1, 16),
                                            (28, 16, 0, 1, 16), (8, 16, 1, 1, 8)]:
            gb.set_tuning(waves, box, xcd, runs, tile)
            gb.frame(True)
            assert frames_equal(gb.read_all(), base) == [], (waves, box, xcd, runs, tile)
    finally:
        gb.finish()


def test_set_tuning_rejects_out_of_range(gpu_batch):
    """srpBatchSetTuning reports values outside waves/CU 1..32, mask box
    0..24 or tile sizes other than 8 and 16 through the message callback and
    keeps the batch as it was."""
    import ctypes
    gb = gpu_batch
    gb.setup(8, 0.05, 13, batched=True)
    CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)

    class Callback(ctypes.Structure):
        _fields_ = [("func", CB), ("userParameter", ctypes.c_void_p)]

    got = []
    cb = CB(lambda kind, sev, fn, text, user: got.append(fn.decode()))
    gb.lib.srpSetMessageCallback.argtypes = [Callback]
    try:
        gb.set_tuning(8, 16, 1, 1)
        gb.frame(True)
        base = gb.read_all()
        gb.lib.srpSetMessageCallback(Callback(cb, None))
        for waves, box, tile in [(0, 16, 8), (33, 16, 8), (8, 25, 8), (8, -1, 8), (8, 16, 12), (8, 16, 32)]:
            gb.set_tuning(waves, box, 1, 1, tile)
        assert got == ["srpBatchSetTuning"] * 6
        gb.frame(True)
        assert frames_equal(gb.read_all(), base) == []
    finally:
        gb.lib.srpSetMessageCallback(Callback(CB(), None))
        gb.finish()


def test_autotuner_picks_a_candidate_and_keeps_results(gpu_batch):
    gb = gpu_batch
    gb.setup(16, 0.05, 9, batched=True)
    try:
        gb.frame(True)
        base = gb.read_all()
        ms, (waves, box, xcd, runs, tile) = gb.autotune(2)
        print(f"[batch] autotune: {ms:.3f} ms with waves/CU {waves}, mask box {box}, xcd queues {xcd}, "
              f"pixel runs {runs}, tile {tile}", flush=True)
        assert ms > 0
        assert waves in (4, 8, 16, 28) and box in (0, 16, 24) and xcd in (0, 1) and runs in (0, 1)
        assert tile in (8, 16)
        gb.frame(True)
        assert frames_equal(gb.read_all(), base) == []
    finally:
        gb.finish()


def _callback(lib, got):
    import ctypes
    CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)

    class Callback(ctypes.Structure):
        _fields_ = [("func", CB), ("userParameter", ctypes.c_void_p)]

    cb = CB(lambda kind, sev, fn, text, user: got.append((fn.decode(), text.decode())))
    lib.srpSetMessageCallback.argtypes = [Callback]
    lib.srpSetMessageCallback(Callback(cb, None))
    return lambda: lib.srpSetMessageCallback(Callback(CB(), None)), cb


@pytest.mark.parametrize("captured", [False, True])
def test_batch_follows_new_buffer_data(gpu_batch, captured):
    """New vertex data for two scenes (their device buffers move) after the
    batch was made, and after its graph was captured: the next frames are
    prepared again and equal drawing the scenes one at a time, with no
    message; the graph is replayed again afterwards."""
    gb = gpu_batch
    gb.setup(12, 0.05, 21, batched=True)
    got = []
    restore, _keep = _callback(gb.lib, got)
    try:
        gb.frame(True)
        if captured:
            assert gb.capture() == 0
            gb.frame(True)
        gb.regenerate(3, 0x1234)
        gb.regenerate(7, 0x5678)
        for r in range(3):
            gb.frame(True)
            batched = gb.read_all()
            gb.frame(False)
            single = gb.read_all()
            assert frames_equal(batched, single) == [], f"frame {r}"
        assert got == []
    finally:
        restore()
        gb.finish()


def test_batch_with_freed_buffer_draws_nothing(gpu_batch):
    """A batch whose vertex buffer was freed reports it through the message
    callback and draws nothing (the framebuffers keep their contents)."""
    gb = gpu_batch
    gb.setup(6, 0.05, 23, batched=True)
    got = []
    restore, _keep = _callback(gb.lib, got)
    try:
        gb.frame(True)
        before = gb.read_all()
        gb.drop_scene(2)
        gb.frame(True)
        gb.frame(True)
        assert [fn for fn, _ in got] == ["srpBatchRender", "srpBatchRender"]
        assert "freed" in got[0][1]
        assert frames_equal(gb.read_all(), before) == []
    finally:
        restore()
        gb.finish()


def test_new_batch_and_capture_draw_nothing(gpu_batch):
    """Making a batch and capturing its graph leave the framebuffers as they
    were: only srpBatchRender draws."""
    gb = gpu_batch
    gb.setup(6, 0.05, 25, batched=False)
    try:
        gb.frame(False)
        ref = gb.read_all()
    finally:
        gb.finish()
    gb.setup(6, 0.05, 25, batched=True)
    try:
        untouched = gb.read_all()
        assert gb.capture() == 0
        assert frames_equal(gb.read_all(), untouched) == []
        gb.frame(True)
        assert frames_equal(gb.read_all(), ref) == []
    finally:
        gb.finish()
