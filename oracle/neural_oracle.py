"""CPU oracle of the neural fragment shader's network (test infrastructure).

Only tests, the smoke check and the benchmark's CPU leg use this file, as
the checker of the GPU kernels (srp_amd/csrc/neural/mlp.hip,
include/srp/device/neural.hpp).  It restates, in float64 numpy:

* the Fourier encoding: index 64*d + 2*j + s holds sin (s=0) or cos (s=1)
  of 2*pi*0.5*(j+1)*x_d, d = 0..7, j = 0..31;
* the network 512 -> 64 -> ReLU -> 64 -> ReLU -> 3 and its loss, the mean
  squared error over the batch's RGB values;
* the backward pass and one Adam step (beta 0.9 / 0.999, eps 1e-8, bias
  corrected);
* the minibatch sampler (splitmix64 finaliser of seed and slot).

`bf16=True` rounds exactly where the GPU hands values to the matrix cores
as bf16 (encoded inputs, weights, hidden activations, pre-activation
gradients); `bf16=False` is the exact float64 function, whose gradient the
tests pin by central differences.  No counterpart exists in the reference
renderer: this network is an extension of the draw path, and its
arithmetic is pinned by closed forms, finite differences and torch.
"""
import numpy as np

N_IN, N_FREQ, N_ENC, N_HID, N_OUT = 8, 32, 512, 64, 3
FREQ_STEP = 0.5
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_ENC
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_OUT * N_HID
N_PARAMS = OFF_B3 + N_OUT
MASK64 = (1 << 64) - 1


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_HID, N_ENC), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(N_OUT, N_HID), p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def encode(F, steps=None):
    """[N, 8] features -> [N, 512] encoding (float64); steps = frequency
    step per feature (default FREQ_STEP for all)."""
    F = np.asarray(F, dtype=np.float64)
    st = np.full(N_IN, FREQ_STEP) if steps is None else np.asarray(steps, dtype=np.float64)
    j = np.arange(N_FREQ, dtype=np.float64) + 1.0
    a = 2.0 * np.pi * (st[None, :] * F)[:, :, None] * j[None, None, :]
    out = np.empty((F.shape[0], N_IN, N_FREQ, 2))
    out[..., 0] = np.sin(a)
    out[..., 1] = np.cos(a)
    return out.reshape(F.shape[0], N_ENC)


def forward(p, F, bf=True, steps=None):
    """Returns (y, cache) with y [N, 3]."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    X = r(encode(F, steps))
    z1 = X @ r(W1).T + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(W2).T + b2
    h2 = r(np.maximum(z2, 0.0))
    y = h2 @ r(W3).T + b3
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(p, F, T, bf=True, steps=None):
    """Mean squared error over N*3 values and its gradient (canonical layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    y, (X, z1, h1, z2, h2) = forward(p, F, bf, steps)
    n = y.shape[0]
    e = y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = r(2.0 * e / (3.0 * n))
    dz2 = r((dy @ r(W3)) * (z2 > 0))
    dz1 = r((dz2 @ r(W2)) * (z1 > 0))
    g = pack(dz1.T @ X, dz1.sum(0), dz2.T @ h1, dz2.sum(0), dy.T @ h2, dy.sum(0))
    return loss, g


def adam(p, m, v, g, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = 1-based step count); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def batch_rows(seed, step, batch, n):
    """Dataset rows of training step `step` (the GPU sampler)."""
    out = np.empty(batch, dtype=np.int64)
    for s in range(batch):
        key = (seed + 0x9E3779B97F4A7C15 * (step * batch + s + 1)) & MASK64
        out[s] = mix64(key) % n
    return out


def init_params(seed):
    """The library's initialisation: uniform in +-1/sqrt(fan_in)."""
    st = seed
    p = np.empty(N_PARAMS)
    b_in = float(np.float32(1.0) / np.sqrt(np.float32(N_ENC)))
    b_hid = float(np.float32(1.0) / np.sqrt(np.float32(N_HID)))
    for i in range(N_PARAMS):
        st = (st + 0x9E3779B97F4A7C15) & MASK64
        u = (mix64(st) >> 11) * (1.0 / 9007199254740992.0)
        p[i] = np.float32((2.0 * u - 1.0) * (b_in if i < OFF_W2 else b_hid))
    return p


def decode_capture(slots):
    """Features from the four capture colour buffers (uint32 R<<24|G<<16|B<<8|A):
    pass k carries features 2k (R,G) and 2k+1 (B,A) as 16-bit values."""
    f = []
    for c in slots:
        c = np.asarray(c, dtype=np.uint32)
        f.append((c >> 16).astype(np.float32) / np.float32(65535.0))
        f.append((c & 0xFFFF).astype(np.float32) / np.float32(65535.0))
    F = np.stack(f, -1)
    F[..., 3:6] = F[..., 3:6] * np.float32(2.0) - np.float32(1.0)
    return F


def psnr(a, b):
    """PSNR (peak 255) and largest difference of two uint8 RGB arrays."""
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return (10.0 * np.log10(255.0 ** 2 / mse) if mse > 0 else float("inf")), float(np.abs(d).max())
