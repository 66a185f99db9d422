"""GPU checks of the neural fragment shader (include/srp/srp_neural.h):
forward pass, gradients and the Adam step against the CPU oracle
(oracle/neural_oracle.py, itself pinned in tests/test_neural_oracle.py);
G-buffer capture against the oracle's frames; and a small end-to-end run in
which the network, trained on the GPU from captured fragments, shades
held-out views of the terrain, with what it gains over no network."""
import ctypes
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))

import neural_oracle as no  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nn():
    from neural_runner import Neural
    return Neural()


def synth(n, seed):
    """Terrain-like features (position and UV in [0,1], upward unit normal)
    and a smooth colour target."""
    rng = np.random.default_rng(seed)
    F = np.empty((n, 8), np.float32)
    F[:, 0:3] = rng.uniform(0, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm[:, 1] = np.abs(nrm[:, 1]) + 1.0
    F[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    F[:, 6] = F[:, 0]
    F[:, 7] = F[:, 2]
    T = np.stack([0.5 + 0.4 * np.sin(6.0 * F[:, 0]) * F[:, 4],
                  0.5 + 0.3 * np.cos(5.0 * F[:, 2] + F[:, 3]),
                  0.4 + 0.3 * F[:, 1] * F[:, 4]], -1).astype(np.float32)
    return F, T


def test_init_matches_oracle(nn):
    net = nn.new_net(3)
    try:
        assert np.array_equal(nn.get_params(net), no.init_params(3).astype(np.float32))
    finally:
        nn.free_net(net)


@pytest.mark.parametrize("scale", [1.0, 4.0])
def test_infer_matches_oracle(nn, scale):
    """Matrix-core inference (64 fragments per wave, ragged tail) and the
    per-lane path against the oracle's bf16-rounded forward pass."""
    n = 4096 + 37
    F, T = synth(n, 1)
    ds = nn.new_ds(8192)
    net = nn.new_net(3)
    try:
        nn.ds_write(ds, F, T)
        p = (no.init_params(3) * scale).astype(np.float32)
        nn.set_params(net, p)
        yw = nn.infer(net, ds, n)
        yl = nn.infer(net, ds, n, lanes=True)
        yo, _ = no.forward(p.astype(np.float64), F, bf=True)
        ew = np.abs(yw - yo)
        el = np.abs(yl - yo)
        print(f"[neural] scale {scale}: wave max {ew.max():.2e} mean {ew.mean():.2e}; "
              f"lane max {el.max():.2e}; |y| {np.abs(yo).mean():.3f}", flush=True)
        # a bf16 rounding flip of one encoded input or hidden value moves an
        # output by up to ~1% of its scale; on average outputs agree far closer
        tol = 1e-2 * max(float(np.abs(yo).mean()), 0.2)
        assert ew.max() < tol and el.max() < tol
        assert ew.mean() < tol / 20
        assert np.abs(yw - yl).max() < tol
    finally:
        nn.free_net(net)
        nn.free_ds(ds)


def test_encoding_steps_match_oracle(nn):
    """Per-feature frequency steps set on the network reach the inference
    paths and the gradient exactly as the oracle's encoding applies them."""
    n = 2048
    F, T = synth(n, 6)
    st = np.array([1.5, 0.5, 0.75, 0.5, 0.5, 0.5, 3.0, 1.25], np.float32)
    ds = nn.new_ds(n)
    net = nn.new_net(8)
    try:
        nn.ds_write(ds, F, T)
        nn.set_encoding(net, st)
        assert np.array_equal(nn.get_encoding(net), st)
        p = (no.init_params(8) * 2.0).astype(np.float32)
        nn.set_params(net, p)
        yo, _ = no.forward(p.astype(np.float64), F, bf=True, steps=st)
        tol = 1e-2 * max(float(np.abs(yo).mean()), 0.2)
        for lanes in (False, True):
            e = np.abs(nn.infer(net, ds, n, lanes=lanes) - yo)
            print(f"[neural] steps, lanes={lanes}: max {e.max():.2e} mean {e.mean():.2e}", flush=True)
            assert e.max() < tol and e.mean() < tol / 20
        g, loss = nn.gradients(net, ds, n)
        lo, go = no.loss_and_grad(p.astype(np.float64), F, T, bf=True, steps=st)
        assert abs(loss - lo) <= 1e-4 * lo
        rel = np.linalg.norm(g - go) / np.linalg.norm(go)
        assert rel < 1e-2
        # a step outside (0, 1024) is refused and changes nothing
        nn.set_encoding(net, np.full(8, -1.0, np.float32))
        assert np.array_equal(nn.get_encoding(net), st)
    finally:
        nn.free_net(net)
        nn.free_ds(ds)


def test_gradients_match_oracle(nn):
    n = 2048
    F, T = synth(n, 2)
    ds = nn.new_ds(n)
    net = nn.new_net(5)
    try:
        nn.ds_write(ds, F, T)
        p = (no.init_params(5) * 3.0).astype(np.float32)
        nn.set_params(net, p)
        g, loss = nn.gradients(net, ds, n)
        lo, go = no.loss_and_grad(p.astype(np.float64), F, T, bf=True)
        assert abs(loss - lo) <= 1e-4 * lo
        blocks = {"W1": (no.OFF_W1, no.OFF_B1), "b1": (no.OFF_B1, no.OFF_W2), "W2": (no.OFF_W2, no.OFF_B2),
                  "b2": (no.OFF_B2, no.OFF_W3), "W3": (no.OFF_W3, no.OFF_B3), "b3": (no.OFF_B3, no.N_PARAMS)}
        for name, (a, b) in blocks.items():
            rel = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            print(f"[neural] grad {name}: rel err {rel:.2e}", flush=True)
            assert rel < 1e-2, name
    finally:
        nn.free_net(net)
        nn.free_ds(ds)


def test_adam_step_matches_oracle(nn):
    """One optimiser step: the GPU's minibatch sampler, loss and Adam update
    equal the oracle's."""
    n, batch, lr, seed = 5000, 1024, 1e-3, 77
    F, T = synth(n, 3)
    ds = nn.new_ds(n)
    net = nn.new_net(9)
    try:
        nn.ds_write(ds, F, T)
        p0 = no.init_params(9)
        loss = nn.train(net, ds, 1, batch, lr, lr, seed)
        p1 = nn.get_params(net).astype(np.float64)
        rows = no.batch_rows(seed, 0, batch, n)
        lo, g = no.loss_and_grad(p0, F[rows], T[rows], bf=True)
        p1o, _, _ = no.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, lr, 1)
        assert abs(float(loss[0]) - lo) <= 1e-4 * lo
        off = np.abs(p1 - p1o) > 0.1 * lr
        print(f"[neural] adam: {off.sum()} of {off.size} parameters differ by > lr/10", flush=True)
        assert off.mean() < 5e-3
    finally:
        nn.free_net(net)
        nn.free_ds(ds)


def test_training_reduces_loss(nn):
    F, T = synth(65536, 4)
    ds = nn.new_ds(65536)
    net = nn.new_net(1)
    try:
        nn.ds_write(ds, F, T)
        loss = nn.train(net, ds, 400, 4096, 3e-3, 3e-4, 5)
        print(f"[neural] loss {loss[:3]} -> {loss[-3:]}", flush=True)
        assert np.all(np.isfinite(loss))
        assert loss[-20:].mean() < loss[:5].mean() / 10
    finally:
        nn.free_net(net)
        nn.free_ds(ds)


def test_capture_matches_oracle_frames(nn, oracle):
    """The dataset built on the GPU from the four capture passes and the
    procedural pass equals the decoding of the oracle's frames."""
    W, H = 480, 270
    assert nn.terrain_setup(W, H, 200, 100, 5, 0) == 0
    ds = nn.new_ds(W * H)
    try:
        got = nn.terrain_collect(ds, 5, 1, 1.0, 0)
        Fg, Tg = nn.ds_read(ds, 0, got)
    finally:
        nn.free_ds(ds)
        nn.finish()
    frames = [oracle.run("cfg/terrain", width=W, height=H, p0=200, p1=100, p2=5, seed=k) for k in range(5)]
    mask = frames[0].depth < 1
    Fo = no.decode_capture([f.color for f in frames[1:]])[mask]
    To = (frames[0].rgb()[mask].astype(np.float32) / np.float32(255.0))
    assert got == mask.sum()
    assert np.array_equal(Fg, Fo)
    assert np.array_equal(Tg, To)


@pytest.mark.parametrize("batch", [0, 1000, 65536 + 1024])
def test_train_rejects_invalid_batch(nn, batch):
    """A batch that is not a positive multiple of 1,024 samples, or above
    65,536, is reported through the message callback; the parameters stay
    as they were."""
    CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)

    class Callback(ctypes.Structure):
        _fields_ = [("func", CB), ("userParameter", ctypes.c_void_p)]

    got = []
    cb = CB(lambda kind, sev, fn, text, user: got.append((fn.decode(), text.decode())))
    nn.lib.srpSetMessageCallback.argtypes = [Callback]
    F, T = synth(4096, 77)
    ds = nn.new_ds(4096)
    net = nn.new_net(5)
    try:
        nn.ds_write(ds, F, T)
        before = nn.get_params(net)
        nn.lib.srpSetMessageCallback(Callback(cb, None))
        nn._train(net, ds, 3, batch, 1e-3, 1e-4, 1, None)
        nn.lib.srpSetMessageCallback(Callback(CB(), None))
        assert len(got) == 1 and got[0][0] == "srpNeuralNetTrain"
        assert f"Batch of {batch} samples" in got[0][1]
        assert np.array_equal(nn.get_params(net), before)
    finally:
        nn.lib.srpSetMessageCallback(Callback(CB(), None))
        nn.free_net(net)
        nn.free_ds(ds)


class _Framebuffer(ctypes.Structure):
    """SRPFramebuffer of include/srp/framebuffer.h."""
    _fields_ = [("width", ctypes.c_size_t), ("height", ctypes.c_size_t), ("size", ctypes.c_size_t),
                ("color", ctypes.POINTER(ctypes.c_uint32)), ("depth", ctypes.POINTER(ctypes.c_float)),
                ("stencil", ctypes.POINTER(ctypes.c_uint8))]


@pytest.mark.parametrize("w,h,covered", [(1920, 1080, 0.6), (7, 3, 0.5), (1, 1, 1.0), (2048, 1031, 0.0)])
def test_compare_matches_numpy(nn, w, h, covered):
    """srpNeuralCompare over whole frames (1080p: each thread strides over
    eight pixels; ragged and empty masks too) equals the PSNR, largest
    difference and pixel count numpy computes over the covered pixels."""
    L = nn.lib
    L.srpNewFramebuffer.restype = ctypes.POINTER(_Framebuffer)
    L.srpNewFramebuffer.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    L.srpFreeFramebuffer.argtypes = [ctypes.POINTER(_Framebuffer)]
    rng = np.random.default_rng(w * 7919 + h)
    fbs = [L.srpNewFramebuffer(w, h) for _ in range(3)]
    try:
        n = w * h
        arrays = []
        for fb in fbs:
            c = np.ctypeslib.as_array(fb.contents.color, (n,))
            d = np.ctypeslib.as_array(fb.contents.depth, (n,))
            c[:] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
            d[:] = np.where(rng.uniform(size=n) < covered, rng.uniform(0, 0.999, n), 1.0).astype(np.float32)
            arrays.append((c.copy(), d.copy()))
        # the second frame differs from the first by a little on most pixels
        ca = arrays[0][0]
        cb = (ca.astype(np.int64) ^ (rng.integers(0, 4, n) << (8 * rng.integers(1, 4, n)))).astype(np.uint32)
        np.ctypeslib.as_array(fbs[1].contents.color, (n,))[:] = cb
        ps, mx, npx = nn.compare(ctypes.cast(fbs[0], ctypes.c_void_p), ctypes.cast(fbs[1], ctypes.c_void_p),
                                 ctypes.cast(fbs[2], ctypes.c_void_p))
        m = arrays[2][1] < 1

        def rgb(c):
            return np.stack([(c >> 24) & 255, (c >> 16) & 255, (c >> 8) & 255], -1).astype(np.uint8)

        assert npx == int(m.sum())
        if npx:
            pn, mn = no.psnr(rgb(ca)[m], rgb(cb)[m])
            assert mx == mn
            assert ps == pytest.approx(pn, abs=1e-9) if np.isfinite(pn) else ps == pn
        else:
            assert mx == 0.0
    finally:
        for fb in fbs:
            L.srpFreeFramebuffer(fb)


@pytest.mark.parametrize("scale", [1.0, 3.0])
def test_deferred_shading_equals_shading_in_the_rasteriser(nn, scale):
    """The network as the terrain's fragment shader, evaluated once per
    visible pixel after the rasteriser (deferred) and for every fragment that
    passes the depth test inside it (the default): the same colour bits,
    depth, stencil and fragment counts, on three poses (each network output
    depends only on its own fragment's inputs, which the resolve rebuilds
    exactly)."""
    W, H = 480, 270
    assert nn.terrain_setup(W, H, 400, 200, 256, 0) == 0
    net = nn.new_net(7)
    assert nn.get_deferred() == 0
    try:
        nn.set_params(net, (no.init_params(7) * scale).astype(np.float32))
        nn.terrain_shader(0, nn.neural_fs, nn.net_device(net))
        for pose in (256, 290, 333):
            print(f"[deferred] scale {scale} pose {pose}", flush=True)
            nn.terrain_pose(pose)
            out = []
            for deferred in (1, 0):
                nn.set_deferred(deferred)
                nn.reset_stats()
                nn.frame()
                out.append((nn.read_frame(W, H), nn.stats()))
            (a, sa), (b, sb) = out
            assert np.array_equal(a[0], b[0]), f"pose {pose}: {np.count_nonzero(a[0] != b[0])} colours differ"
            assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
            assert np.array_equal(a[2], b[2])
            assert sa == sb, (sa, sb)
            assert np.count_nonzero(a[1] < 1) > W * H // 10   # the terrain covers the frame
    finally:
        nn.set_deferred(0)
        nn.free_net(net)
        nn.finish()


def test_neural_shader_end_to_end(nn):
    """Capture -> GPU training -> the network as the terrain's fragment
    shader on held-out poses.  Coverage, depth and stencil equal the
    procedural frames exactly; colour is judged by PSNR against them, next
    to an untrained network and the best constant colour (no network)."""
    W, H = 480, 270
    assert nn.terrain_setup(W, H, 1000, 500, 0, 0) == 0
    ds = nn.new_ds(2_000_000)
    net = nn.new_net(21)
    try:
        got = nn.terrain_collect(ds, 0, 48, 0.35, 0x5EED0005)
        print(f"[neural] {got} training samples", flush=True)
        _, Tall = nn.ds_read(ds, 0, got)
        mean = Tall.mean(0)
        dev = nn.net_device(net)

        def evaluate():
            res = []
            for pose in range(256, 262):
                nn.terrain_pose(pose)
                nn.use_fb(0)
                nn.terrain_shader(0, None, None)
                nn.frame()
                fa = nn.fb()
                plain = nn.read_frame(W, H)
                nn.use_fb(1)
                nn.terrain_shader(0, nn.neural_fs, dev)
                nn.frame()
                fb = nn.fb()
                neural = nn.read_frame(W, H)
                ps, mx, npx = nn.compare(fa, fb, fa)
                assert np.array_equal(plain[1].view(np.uint32), neural[1].view(np.uint32))
                assert np.array_equal(plain[2], neural[2])
                m = plain[1] < 1
                const = np.broadcast_to(np.round(mean * 255).astype(np.uint8), (int(m.sum()), 3))
                pc, _ = no.psnr(const, no_rgb(plain[0])[m])
                pn, mn = no.psnr(no_rgb(neural[0])[m], no_rgb(plain[0])[m])
                assert abs(pn - ps) < 1e-6 and mn == mx and npx == m.sum()
                res.append((ps, mx, pc))
            return np.array(res)

        before = evaluate()
        loss = nn.train(net, ds, 3000, 8192, 3e-3, 1e-4, 17)
        after = evaluate()
        print(f"[neural] loss {loss[0]:.4f} -> {loss[-1]:.6f}; PSNR untrained {before[:, 0].mean():.2f} dB, "
              f"constant colour {after[:, 2].mean():.2f} dB, trained {after[:, 0].mean():.2f} dB "
              f"(max err {after[:, 1].max():.0f})", flush=True)
        assert after[:, 0].mean() > 30.0
        assert after[:, 0].mean() - before[:, 0].mean() > 15.0
        assert after[:, 0].mean() - after[:, 2].mean() > 10.0
    finally:
        nn.use_fb(0)
        nn.free_net(net)
        nn.free_ds(ds)
        nn.finish()


def no_rgb(color):
    from neural_runner import rgb_of
    return rgb_of(color)
