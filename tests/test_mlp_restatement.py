"""CPU checks of the surrogate's numpy restatement (tests/mlp_ref.py):
the backward pass against central finite differences, the feature
identities, the BF16 / E4M3 rounding, the draw sequence, and the
parameter layout the C ABI reports."""
import numpy as np
import pytest

import mlp_ref as R


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for i, o in R.DIMS:
        out.append(rng.uniform(-1, 1, (o, i)).reshape(-1) * np.sqrt(6.0 / i) * scale)
        out.append(rng.normal(0, 0.05, o))
    return np.concatenate(out)


def test_param_count_matches_library():
    from bqp_amd import _native, surrogate
    assert R.N_PARAMS == 202_499 == surrogate.N_PARAMS
    assert _native.load().bqp_surrogate_param_count() == R.N_PARAMS


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    X = rng.uniform(-1, 1, (6, 16))
    T = rng.normal(0, 1, (6, 3))
    L, g = R.loss_grad(p, X, T)
    assert np.isclose(L, R.loss(p, X, T))
    # probe weights and biases of every layer
    offs, off = [], 0
    for i, o in R.DIMS:
        offs += [off + int(k) for k in rng.integers(0, i * o, 6)]
        off += i * o
        offs += [off + int(k) for k in rng.integers(0, o, 3)]
        off += o
    h = 1e-6
    for k in offs:
        pp = p.copy()
        pm = p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss(pp, X, T) - R.loss(pm, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_backward_detects_a_planted_sign_error():
    # the finite-difference check above would catch a dsilu sign slip
    rng = np.random.default_rng(2)
    p = _params(2)
    X = rng.uniform(-1, 1, (4, 16))
    T = rng.normal(0, 1, (4, 3))
    good = R.dsilu
    try:
        R.dsilu = lambda z: (1.0 / (1.0 + np.exp(-z))) * (1.0 - z * (1.0 - 1.0 / (1.0 + np.exp(-z))))
        _, g = R.loss_grad(p, X, T)
    finally:
        R.dsilu = good
    k = 5  # a first-layer weight: its gradient passes through every dsilu
    pp, pm = p.copy(), p.copy()
    pp[k] += 1e-6
    pm[k] -= 1e-6
    fd = (R.loss(pp, X, T) - R.loss(pm, X, T)) / 2e-6
    assert abs(fd - g[k]) > 1e-6 * max(1.0, abs(fd))


def test_feature_identities():
    rng = np.random.default_rng(3)
    d = rng.normal(size=(100, 3))
    pos = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(6578e3, 8378e3, 100)[:, None]
    f = R.features(pos, 6578e3, 8378e3)
    assert np.allclose(np.linalg.norm(f[:, :3], axis=1), 1.0)
    assert np.all(np.abs(f[:, 3]) <= 1.0)
    for i in range(3):
        s1, c1, s2, c2 = f[:, 4 + 2 * i], f[:, 5 + 2 * i], f[:, 10 + 2 * i], f[:, 11 + 2 * i]
        assert np.allclose(s1 ** 2 + c1 ** 2, 1.0)
        assert np.allclose(s2, 2 * s1 * c1)
        assert np.allclose(c2, 1 - 2 * s1 ** 2)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 65504.0], dtype=np.float32)
    y = R.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0  # tie -> even
    assert y[2] == 1.0 + 4 * 2 ** -8  # tie -> even (up)
    assert y[3] == 1.0
    assert abs(y[4] - x[4]) <= 2 ** -8 * abs(x[4])
    # bf16 has 8 significant bits: rounding again is the identity
    assert np.array_equal(R.bf16(y), y)


def test_e4m3_codes():
    codes = np.arange(0, 0x7F, dtype=np.int64)
    vals = R.e4m3_decode(codes)
    assert vals[0] == 0.0 and vals[-1] == 448.0 and vals[1] == 2.0 ** -9 and vals[8] == 2.0 ** -6
    assert np.all(np.diff(vals) > 0)
    # every finite value encodes to its own code, both signs
    assert np.array_equal(R.e4m3_encode(vals), codes.astype(np.uint8))
    assert np.array_equal(R.e4m3_encode(-vals[1:]), (codes[1:] | 0x80).astype(np.uint8))
    # midpoints round to the even code; beyond the range saturates
    mids = 0.5 * (vals[:-1] + vals[1:])
    enc = R.e4m3_encode(mids).astype(np.int64)
    assert np.all(enc % 2 == 0)
    assert np.all((enc == codes[:-1]) | (enc == codes[1:]))
    assert R.e4m3_encode(1e6) == 0x7E and R.e4m3_encode(-500.0) == 0xFE


def test_minibatch_draws_cover_rows_uniformly():
    rows = R.minibatch_rows(7, 0, 200_000, 10)
    counts = np.bincount(rows, minlength=10)
    assert rows.min() >= 0 and rows.max() < 10
    assert np.all(np.abs(counts - 20_000) < 5 * np.sqrt(20_000))
    assert not np.array_equal(R.minibatch_rows(7, 1, 100, 1000), R.minibatch_rows(7, 0, 100, 1000))
    # splitmix64 known answer: first output of the generator seeded with 0
    assert int(R.mix64(np.uint64(0))) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_quantised_inference_close_to_fp64(fmt):
    rng = np.random.default_rng(4)
    p = _params(4, 0.5)
    X = rng.uniform(-1, 1, (64, 16))
    Y, _, _ = R.forward(p, X)
    Yq = R.infer_bf16(p, X) if fmt == "bf16" else R.infer_fp8(p, X, X)
    rel = np.sqrt(np.mean((Yq - Y) ** 2) / np.mean(Y ** 2))
    assert rel < (0.02 if fmt == "bf16" else 0.15), rel
