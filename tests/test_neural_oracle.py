"""CPU checks of the neural shader's oracle (oracle/neural_oracle.py).

Each part of the oracle is pinned by something other than its own formula:
the gradient by central differences, the forward pass and Adam by torch,
bf16 rounding by torch's conversion, the encoding by closed forms, the
capture decoding by an invariant of the terrain (UV is an affine function
of the position) on frames the C oracle renders.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
sys.path.insert(0, HERE)

import neural_oracle as no  # noqa: E402


def _data(n, seed):
    rng = np.random.default_rng(seed)
    F = rng.uniform(0, 1, (n, 8))
    F[:, 3:6] = rng.uniform(-1, 1, (n, 3))
    T = rng.uniform(0, 1, (n, 3))
    return F, T


def test_gradient_matches_central_differences():
    p = no.init_params(11)
    F, T = _data(48, 1)
    _, g = no.loss_and_grad(p, F, T, bf=False)
    rng = np.random.default_rng(2)
    idx = list(rng.choice(no.N_PARAMS, 60, replace=False))
    idx += [no.OFF_B1 + 7, no.OFF_B2 + 9, no.OFF_B3 + 2, no.OFF_W3 + 70, no.OFF_W2 + 100]
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = no.loss_and_grad(q, F, T, bf=False)
        q[i] -= 2 * h
        lm, _ = no.loss_and_grad(q, F, T, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(g[i]), 1e-3), (i, fd, g[i])


def test_forward_and_gradient_match_torch():
    torch = pytest.importorskip("torch")
    p = no.init_params(5)
    F, T = _data(32, 3)
    W1, b1, W2, b2, W3, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in no.unpack(p))
    X = torch.tensor(no.encode(F))
    # the encoding by its definition, through torch's own sin/cos
    j = torch.arange(32, dtype=torch.float64) + 1
    a = 2 * np.pi * 0.5 * torch.tensor(F)[:, :, None] * j
    Xt = torch.stack([torch.sin(a), torch.cos(a)], -1).reshape(32, 512)
    assert torch.allclose(X, Xt, atol=1e-12)
    y = torch.relu(torch.relu(Xt @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    loss = ((y - torch.tensor(T)) ** 2).mean()
    loss.backward()
    yo, _ = no.forward(p, F, bf=False)
    lo, g = no.loss_and_grad(p, F, T, bf=False)
    assert np.allclose(yo, y.detach().numpy(), atol=1e-12)
    assert abs(lo - loss.item()) < 1e-12
    gt = no.pack(*(t.grad.numpy() for t in (W1, b1, W2, b2, W3, b3)))
    assert np.allclose(g, gt, atol=1e-12)


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    p = rng.normal(size=100)
    w = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(100)
    v = np.zeros(100)
    for t in range(1, 6):
        g = rng.normal(size=100)
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = no.adam(p, m, v, g, 3e-3, t)
    assert np.allclose(p, w.detach().numpy(), atol=1e-13)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, 5000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 3.0e38, 1e-40]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(no.bf16(x), ref)


def test_encoding_closed_forms():
    F = np.zeros((3, 8))
    F[1, :] = 1.0     # 2*pi*0.5*(j+1)*1 = pi*(j+1): sin 0, cos (-1)^(j+1)
    F[2, :] = 0.5     # pi*(j+1)/2
    E = no.encode(F).reshape(3, 8, 32, 2)
    j = np.arange(32) + 1
    assert np.allclose(E[0, :, :, 0], 0) and np.allclose(E[0, :, :, 1], 1)
    assert np.allclose(E[1, :, :, 0], 0, atol=1e-12)
    assert np.allclose(E[1, :, :, 1], (-1.0) ** j)
    assert np.allclose(E[2, :, :, 0], np.sin(np.pi * j / 2))
    assert np.allclose(E[:, :, :, 0] ** 2 + E[:, :, :, 1] ** 2, 1.0)
    # odd / even symmetry in x
    Fr = np.random.default_rng(7).uniform(-1, 1, (10, 8))
    Ep = no.encode(Fr).reshape(10, 8, 32, 2)
    Em = no.encode(-Fr).reshape(10, 8, 32, 2)
    assert np.allclose(Ep[..., 0], -Em[..., 0]) and np.allclose(Ep[..., 1], Em[..., 1])


def test_encoding_per_feature_steps():
    """A frequency step per feature scales that feature alone: encoding with
    steps s equals the default encoding of the features times s/FREQ_STEP,
    and a quarter-revolution input gives the closed form per feature."""
    rng = np.random.default_rng(11)
    F = rng.uniform(-1, 1, (6, 8))
    st = np.array([0.5, 1.5, 0.75, 0.25, 2.0, 1.0, 3.0, 0.5])
    assert np.allclose(no.encode(F, st), no.encode(F * st / no.FREQ_STEP), atol=1e-12)
    E = no.encode(np.tile(0.25 / st, (1, 1)), st).reshape(1, 8, 32, 2)
    j = np.arange(32) + 1
    for d in range(8):
        assert np.allclose(E[0, d, :, 0], np.sin(np.pi * j / 2), atol=1e-12)
        assert np.allclose(E[0, d, :, 1], np.cos(np.pi * j / 2), atol=1e-12)
    # a different step on one feature changes only that feature's block
    st2 = st.copy()
    st2[3] = 0.6
    d = np.abs(no.encode(F, st) - no.encode(F, st2)).reshape(6, 8, 64).max(axis=(0, 2))
    assert d[3] > 1e-3 and np.all(np.delete(d, 3) == 0)


def test_sampler_is_deterministic_and_in_range():
    a = no.batch_rows(9, 3, 256, 1000)
    b = no.batch_rows(9, 3, 256, 1000)
    c = no.batch_rows(9, 4, 256, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() < 1000
    # roughly uniform
    h = np.bincount(no.batch_rows(1, 0, 20000, 10), minlength=10)
    assert h.min() > 1700 and h.max() < 2300


def test_capture_decoding_on_oracle_frames():
    """The capture passes of the C oracle decode to features whose UV equals
    the normalised x/z position (the terrain's UV is affine in position) and
    whose normal is unit length up to interpolation."""
    from scene_runner import ORACLE_SO, SceneModule
    if not os.path.exists(ORACLE_SO):
        pytest.skip("oracle module not built")
    m = SceneModule(ORACLE_SO)
    caps = [m.run("cfg/terrain", width=240, height=135, p0=200, p1=100, p2=3, seed=1 + k) for k in range(4)]
    mask = caps[0].depth < 1
    assert mask.sum() > 5000
    F = no.decode_capture([c.color for c in caps])[mask]
    q = 2.0 / 65535
    assert np.abs(F[:, 6] - F[:, 0]).max() < q
    assert np.abs(F[:, 7] - F[:, 2]).max() < q
    n = np.linalg.norm(F[:, 3:6], axis=1)
    assert np.all(n > 0.9) and np.all(n < 1.0 + 4 * q)
    assert F[:, 4].min() > 0          # normals point up
    assert np.all((F[:, :3] >= 0) & (F[:, :3] <= 1))
