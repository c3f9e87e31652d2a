"""FP64 numpy restatement of the surrogate network (test infrastructure).

Forward, loss and backward of  16 -> 4 x [256, SiLU] -> 3  with the flat
parameter layout of the library (W_l row-major (out x in), then b_l), the
16 position features, BF16 / FP8 E4M3 rounding, and the splitmix64 draws of
the minibatch sampler.  The backward pass here is checked against central
finite differences (tests/test_mlp_restatement.py), so it can serve as the
reference for the device gradients.
"""
from __future__ import annotations

import numpy as np

DIMS = [(16, 256), (256, 256), (256, 256), (256, 256), (256, 3)]  # (in, out)
N_PARAMS = sum(i * o + o for i, o in DIMS)

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for i, o in DIMS:
        W = p[off:off + i * o].reshape(o, i)
        off += i * o
        b = p[off:off + o]
        off += o
        out.append((W, b))
    assert off == p.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, X):
    layers = unpack(params)
    h = np.asarray(X, dtype=np.float64)
    Zs, Hs = [], [h]
    for l, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if l < len(layers) - 1:
            Zs.append(z)
            h = silu(z)
            Hs.append(h)
        else:
            return z, Zs, Hs


def loss(params, X, T):
    Y, _, _ = forward(params, X)
    return float(np.mean((Y - np.asarray(T, dtype=np.float64)) ** 2))


def loss_grad(params, X, T):
    """Mean squared error over rows and the 3 outputs, and its gradient."""
    layers = unpack(params)
    Y, Zs, Hs = forward(params, X)
    T = np.asarray(T, dtype=np.float64)
    B = Y.shape[0]
    L = float(np.mean((Y - T) ** 2))
    dZ = 2.0 * (Y - T) / (3.0 * B)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        hprev = Hs[l]
        grads[l] = (dZ.T @ hprev, dZ.sum(axis=0))
        if l > 0:
            dZ = (dZ @ W) * dsilu(Zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])
    return L, flat


def features(pos, r_lo, r_hi):
    """The 16 features of body-fixed positions (n, 3), in FP64."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    r = np.linalg.norm(pos, axis=1)
    u = pos / r[:, None]
    s = (r - 0.5 * (r_lo + r_hi)) / (0.5 * (r_hi - r_lo))
    f = np.empty((pos.shape[0], 16))
    f[:, 0:3] = u
    f[:, 3] = s
    for i in range(3):
        f[:, 4 + 2 * i] = np.sin(np.pi * u[:, i])
        f[:, 5 + 2 * i] = np.cos(np.pi * u[:, i])
        f[:, 10 + 2 * i] = np.sin(2 * np.pi * u[:, i])
        f[:, 11 + 2 * i] = np.cos(2 * np.pi * u[:, i])
    return f


# ------------------------------------------------------------ rounding
def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def e4m3_decode(code):
    c = np.asarray(code, dtype=np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 15
    m = c & 7
    v = np.where(e == 0, np.ldexp(m / 8.0, -6), np.ldexp(1.0 + m / 8.0, e - 7))
    v = np.where((e == 15) & (m == 7), np.nan, v)
    return s * v


_E4M3_POS = np.array([e4m3_decode(c) for c in range(0x7F)])  # codes 0..126, increasing


def e4m3_encode(x):
    """float -> OCP E4M3 code, round to nearest (ties to even code),
    saturating at 448."""
    x = np.asarray(x, dtype=np.float64)
    a = np.minimum(np.abs(x), 448.0)
    idx = np.searchsorted(_E4M3_POS, a)  # first value >= a
    idx = np.clip(idx, 0, 126)
    lo = np.clip(idx - 1, 0, 126)
    dlo = a - _E4M3_POS[lo]
    dhi = _E4M3_POS[idx] - a
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (idx % 2 == 0))
    code = np.where(pick_hi, idx, lo)
    code = np.where(a == _E4M3_POS[idx], idx, code)
    return (code | np.where(np.signbit(x), 0x80, 0)).astype(np.uint8)


def e4m3(x):
    return e4m3_decode(e4m3_encode(x))


# -------------------------------------------------------------- draws
def mix64(z):
    with np.errstate(over="ignore"):  # arithmetic modulo 2^64
        z = (np.asarray(z, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def minibatch_rows(seed: int, step: int, batch: int, n: int):
    """Row indices of training step `step` (with replacement)."""
    sd = mix64(np.uint64(seed) ^ mix64(np.uint64(0x5EED + step)))
    i = np.arange(batch, dtype=np.uint64)
    return (mix64(sd ^ mix64(i)) % np.uint64(n)).astype(np.int64)


# --------------------------------------------------------- emulation
def infer_bf16(params, X):
    """Inference with BF16 weights and layer inputs, FP32-accurate sums."""
    layers = unpack(params)
    h = bf16(np.asarray(X, dtype=np.float32)).astype(np.float64)
    for l, (W, b) in enumerate(layers):
        z = (h @ bf16(W.astype(np.float32)).astype(np.float64).T + b).astype(np.float32)
        if l < len(layers) - 1:
            h = bf16(silu(z.astype(np.float64)).astype(np.float32)).astype(np.float64)
        else:
            return z.astype(np.float64)


def infer_fp8(params, X, calib):
    """Inference with the three 256 x 256 layers in FP8 E4M3: weight scale
    max|W|/448, activation scale max|h|/448 over `calib` rows (FP64 forward),
    layer 0 and the output layer in BF16."""
    layers = unpack(params)
    _, _, Hc = forward(params, calib)
    amax = [np.max(np.abs(Hc[l + 1])) for l in range(3)]  # inputs of hidden layers 1..3
    h = bf16(np.asarray(X, dtype=np.float32)).astype(np.float64)
    W0, b0 = layers[0]
    z = (h @ bf16(W0.astype(np.float32)).astype(np.float64).T + b0).astype(np.float32)
    h = silu(z.astype(np.float64))
    for k in range(3):
        W, b = layers[k + 1]
        sw = np.float32(np.max(np.abs(W.astype(np.float32))) / np.float32(448.0))
        sa = np.float32(np.float32(amax[k]) / np.float32(448.0))
        wq = e4m3(W.astype(np.float32) / sw)
        hq = e4m3((h.astype(np.float32) * np.float32(1.0 / sa)).astype(np.float32))
        z = ((hq @ wq.T) * float(np.float32(sw * sa)) + b).astype(np.float32)
        h = silu(z.astype(np.float64))
    W4, b4 = layers[4]
    h = bf16(h.astype(np.float32)).astype(np.float64)
    return (h @ bf16(W4.astype(np.float32)).astype(np.float64).T + b4).astype(np.float32).astype(np.float64)
