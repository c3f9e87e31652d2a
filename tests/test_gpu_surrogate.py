"""GPU tests of the learned gravity surrogate (device data, training,
MFMA inference, in-integrator fallback) against the CPU oracle and the FP64
numpy restatement in tests/mlp_ref.py.

Bars: training data equal to the oracle residual to FP32 rounding; device
gradients equal to the FP64 restatement to FP32 accumulation error; one
Adam step equal to the restated update; BF16 / FP8 inference equal to their
numpy emulations up to rounding flips; the exact-field fallback bit-identical
to the FP64 path; and the network beating the same run without it.
"""
import numpy as np
import pytest

import bqp_amd
import mlp_ref as R
from bqp_amd import inputs
from gpu_util import announce, pos_err

pytestmark = pytest.mark.gpu

R_LO, R_HI = 6578.0e3, 8378.0e3


@pytest.fixture(scope="module")
def field20():
    return bqp_amd.kaula(20)


def _surrogate(field, seed=0, **kw):
    prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-10, **kw)
    sur = bqp_amd.Surrogate(prop, nlow=4, r_lo=R_LO, r_hi=R_HI, seed=seed)
    return prop, sur


def test_dataset_matches_oracle_residual(orc, field20):
    announce("surrogate dataset: 4096 rows, N=20 residual above degree 4 vs oracle")
    prop, sur = _surrogate(field20)
    feat, targ, pos = sur.dataset(4096, seed=3, positions=True)
    P, F, T = pos.numpy(), feat.numpy(), targ.numpy()
    r = np.linalg.norm(P, axis=1)
    assert r.min() >= R_LO and r.max() <= R_HI
    assert np.abs(F - R.features(P, R_LO, R_HI)).max() < 2e-5
    f = field20
    full = orc.Field(f.cbar, f.sbar, f.nmax, f.mu, f.radius)
    lo = f.truncated(4)
    low = orc.Field(lo.cbar, lo.sbar, 4, f.mu, f.radius)
    res = full.accel(P) - low.accel(P)
    scale = sur.out_scale
    err = np.abs(T * scale - res)
    assert err.max() <= 1e-6 * np.abs(res).max(), err.max()
    # the output scales are the per-component RMS of the residual
    assert np.allclose(np.sqrt(np.mean(T ** 2, axis=0)), 1.0, atol=0.05)
    prop.close()


def test_gradient_matches_fp64_restatement(field20):
    announce("surrogate gradient: 64 rows vs FP64 numpy restatement")
    prop, sur = _surrogate(field20, seed=5)
    feat, targ, _ = sur.dataset(64, seed=9)
    loss, g = sur.loss_grad(feat, targ)
    p = sur.params().astype(np.float64)
    L, gref = R.loss_grad(p, feat.numpy(), targ.numpy())
    assert abs(loss - L) <= 1e-5 * L
    off = 0
    for i, o in R.DIMS:
        for n in (i * o, o):
            a, b = g[off:off + n], gref[off:off + n]
            assert np.abs(a - b).max() <= 2e-4 * np.abs(b).max() + 1e-12, (off, np.abs(a - b).max())
            off += n
    prop.close()


def test_one_adam_step_matches_restatement(field20):
    announce("surrogate Adam: one step, batch 256, vs restated update")
    prop, sur = _surrogate(field20, seed=6)
    feat, targ, _ = sur.dataset(1000, seed=2)
    p0 = sur.params().astype(np.float64)
    lr = 1e-3
    sur.train(feat, targ, steps=1, batch=256, lr=lr, seed=11)
    p1 = sur.params().astype(np.float64)
    rows = R.minibatch_rows(11, 0, 256, 1000)
    _, g = R.loss_grad(p0, feat.numpy()[rows], targ.numpy()[rows])
    # first Adam step: m = 0.1 g, v = 0.001 g^2 -> update lr g / (|g| + 1e-8)
    want = -lr * g / (np.abs(g) + 1e-8)
    big = np.abs(g) > 1e-5
    assert np.mean(np.abs((p1 - p0)[big] - want[big]) <= 1e-3 * lr) > 0.999
    assert np.abs(p1 - p0).max() <= 1.0001 * lr
    prop.close()


@pytest.fixture(scope="module")
def trained(field20):
    announce("surrogate training: 262,144 rows, 1500 steps x 4096 (N=20)")
    prop, sur = _surrogate(field20, seed=1)
    feat, targ, _ = sur.dataset(262144, seed=21)
    losses = sur.fit(feat, targ, steps=1500, batch=4096, lr_max=2e-3, lr_min=2e-5)
    yield prop, sur, feat, targ, losses
    prop.close()


def test_training_reduces_loss(trained):
    _, _, _, _, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-50:].mean() < 0.1 * losses[:10].mean(), (losses[:10].mean(), losses[-50:].mean())


def test_fp32_inference_matches_restatement(trained):
    _, sur, feat, _, _ = trained
    X = feat.rows(0, 2048)
    out, _ = sur.infer(X, "fp32")
    Y, _, _ = R.forward(sur.params().astype(np.float64), X.numpy())
    assert np.abs(out.numpy() - Y).max() <= 1e-4 * np.abs(Y).max()


def test_bf16_inference_matches_emulation(trained):
    announce("surrogate BF16 MFMA inference vs emulation")
    _, sur, feat, _, _ = trained
    sur.pack("bf16")
    X = feat.rows(0, 4096)
    out, ms = sur.infer(X, "bf16")
    Yg = out.numpy()
    Ye = R.infer_bf16(sur.params(), X.numpy())
    d = Yg - Ye
    assert np.sqrt(np.mean(d ** 2)) <= 2e-3 * np.sqrt(np.mean(Ye ** 2)), np.sqrt(np.mean(d ** 2))
    assert np.abs(d).max() <= 2e-2 * np.abs(Ye).max()
    # ragged tail (not a multiple of the 32-row wave tile) gives the same rows
    out2, _ = sur.infer(feat.rows(0, 1000), "bf16")
    assert np.array_equal(out2.numpy(), Yg[:1000])


def test_fp8_inference_matches_emulation(trained):
    announce("surrogate FP8 MFMA inference vs emulation")
    _, sur, feat, _, _ = trained
    calib = feat.rows(0, 8192)
    sur.pack("fp8", calib=calib)
    X = feat.rows(0, 4096)
    out, _ = sur.infer(X, "fp8")
    Yg = out.numpy()
    Ye = R.infer_fp8(sur.params(), X.numpy(), calib.numpy())
    Y64, _, _ = R.forward(sur.params().astype(np.float64), X.numpy())
    rms = np.sqrt(np.mean(Y64 ** 2))
    assert np.sqrt(np.mean((Yg - Ye) ** 2)) <= 0.02 * rms
    assert np.sqrt(np.mean((Yg - Y64) ** 2)) <= 0.2 * rms
    sur.pack("bf16")


def _leo(n, field_n, span):
    w = inputs.workload("w1", n=n, seed=4)
    return w["y0"], np.zeros(n), np.full(n, span)


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16])
def test_failed_check_falls_back_bit_identically(field20, G):
    announce(f"surrogate fallback G={G}: threshold 0 -> every row exact; vs FP64 bitwise")
    y0, t0, tf = _leo(64, 20, 7200.0)
    with bqp_amd.Propagator(field20, method="rk87", rtol=1e-10, launch={"waves_per_group": G}) as p:
        yref, sref, iref = p.propagate_batch_final(y0, t0, tf)
    prop, sur = _surrogate(field20, seed=2, precision="surrogate", launch={"waves_per_group": G})
    sur.pack("bf16" if G != 2 else "fp8", calib=sur.dataset(4096, seed=5)[0])
    sur.set_check(1, 0.0)
    y, s, info = prop.propagate_batch_final(y0, t0, tf)
    st = sur.last_stats()
    assert np.all(s == 0) and np.all(sref == 0)
    assert np.array_equal(y, yref)
    assert st["check_failures"] == 64 and st["net_evals"] == 0 and st["exact_evals"] == info["evals"]
    prop.close()


def test_outside_band_uses_exact_field(field20):
    announce("surrogate band: orbits outside [r_lo, r_hi] -> exact; vs FP64 G=1 bitwise")
    y0, t0, tf = _leo(64, 20, 3600.0)
    with bqp_amd.Propagator(field20, method="rk87", rtol=1e-10, launch={"waves_per_group": 1}) as p:
        yref, _, _ = p.propagate_batch_final(y0, t0, tf)
    prop = bqp_amd.Propagator(field20, method="rk87", rtol=1e-10, precision="surrogate",
                              launch={"waves_per_group": 1})
    sur = bqp_amd.Surrogate(prop, nlow=4, r_lo=20000e3, r_hi=30000e3, seed=2)
    sur.pack("bf16")
    sur.set_check(0, 1e-9)
    y, s, info = prop.propagate_batch_final(y0, t0, tf)
    st = sur.last_stats()
    assert np.array_equal(y, yref)
    assert st["band_fallbacks"] == info["evals"] == st["exact_evals"] and st["net_evals"] == 0
    prop.close()


def test_network_beats_truncated_field(trained, field20):
    """Ablation: with the network (no checks) vs without it (the exact
    degree <= 4 part alone), both against the exact degree-20 field."""
    announce("surrogate ablation: 64 LEO rows x 1 orbit, network vs truncated field")
    prop, sur, _, _, _ = trained
    y0, t0, tf = _leo(64, 20, 6000.0)
    with bqp_amd.Propagator(field20, method="rk87", rtol=1e-10, launch={"waves_per_group": 1}) as p:
        yref, _, _ = p.propagate_batch_final(y0, t0, tf)
    with bqp_amd.Propagator(field20.truncated(4), method="rk87", rtol=1e-10) as p:
        ytr, _, _ = p.propagate_batch_final(y0, t0, tf)
    sur.pack("bf16")
    sur.set_check(0, 1e-9)
    prop.reconfigure(precision="surrogate")
    y, s, info = prop.propagate_batch_final(y0, t0, tf)
    st = sur.last_stats()
    prop.reconfigure(precision="fp64")
    assert np.all(s == 0)
    assert st["net_evals"] == info["evals"]
    e_net, e_tr = pos_err(y, yref), pos_err(ytr, yref)
    announce(f"median position error after 6000 s: network {np.median(e_net):.3g} m, "
             f"truncated field {np.median(e_tr):.3g} m")
    assert np.median(e_net) < 0.5 * np.median(e_tr)
    # the residual network split over the waves of a group gives the same
    # network values, so the same trajectories
    prop.reconfigure(precision="surrogate")
    prop.set_launch(waves_per_group=4)
    y4, s4, info4 = prop.propagate_batch_final(y0, t0, tf)
    prop.set_launch(waves_per_group=0)
    prop.reconfigure(precision="fp64")
    assert np.all(s4 == 0)
    assert np.abs(pos_err(y4, y)).max() < 1e-3 * np.median(e_net) + 1e-6


def test_upfront_first_check_matches_in_kernel_checks(trained, field20):
    """The first sampled check is made for all rows before the propagation
    and failing rows run the plain FP64 kernel; with the same threshold the
    in-kernel path (taken under HIP-graph replay) gives identical rows."""
    announce("surrogate up-front check vs in-kernel checks: 64 rows x 1 h, threshold at the median error")
    prop, sur, _, _, _ = trained
    y0, t0, tf = _leo(64, 20, 3600.0)
    pos = y0[:, :3]  # body frame = inertial at t = 0 (theta0 = 0)
    full = prop.gravity_accel(pos)
    with bqp_amd.Propagator(field20.truncated(4)) as p4:
        low = p4.gravity_accel(pos)
    sur.pack("bf16")
    from bqp_amd.device import DeviceArray
    f = DeviceArray.from_numpy(R.features(pos, R_LO, R_HI).astype(np.float32))
    net, _ = sur.infer(f, "bf16")
    err = np.linalg.norm(low + net.numpy() * sur.out_scale - full, axis=1)
    thr = float(np.median(err))
    sur.set_check(1_000_000_000, thr)  # only evaluation 0 is a check: the up-front split decides
    prop.reconfigure(precision="surrogate")
    prop.set_launch(waves_per_group=4, use_graph=0)
    ya, sa, ia = prop.propagate_batch_final(y0, t0, tf)
    st_a = sur.last_stats()
    prop.set_launch(waves_per_group=4, use_graph=1)
    yb, sb, ib = prop.propagate_batch_final(y0, t0, tf)
    st_b = sur.last_stats()
    prop.set_launch(waves_per_group=0, use_graph=0)
    prop.reconfigure(precision="fp64")
    assert np.all(sa == 0) and np.all(sb == 0)
    assert 8 <= st_a["check_failures"] <= 56 and st_a["checks"] == 64
    assert st_a["net_evals"] > 0
    assert np.array_equal(ya, yb)
    assert st_a == st_b and ia["evals"] == ib["evals"]
