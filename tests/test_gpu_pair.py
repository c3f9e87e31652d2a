"""GPU tests of the pair kernel (bqp_team.hpp): two 64-trajectory groups per
8-wave workgroup, each group's gravity in the eight column shares of the
fixed-G kernel at G = 8, so that a group whose partner is done runs on all
eight waves.  Taken by FP64, split and Encke calls with waves_per_group = 8
and teams_per_block = 2, and by the automatic launch where it measured
faster.

* Every result -- final states, step, evaluation and reject counts, status
  -- equals the fixed-G kernel's at G = 8 bit for bit: RK8(7)13M with
  Vallado drag on ragged durations with more rows than the dealt round
  (queue refills), DOPRI5 (FSAL reuse), DOP853, the compat controller,
  split precision, Encke, sampled output, failing rows, and a batch smaller
  than one pair.
* against the oracle to the usual bars.
"""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from gpu_util import announce, oracle_pair, pos_err

pytestmark = pytest.mark.gpu

DAY = 86400.0
PAIR = {"waves_per_group": 8, "teams_per_block": 2}
FIXED8 = {"waves_per_group": 8, "teams_per_block": 0}


def _both(field, y0, t0, tf, bc=None, **kw):
    out = []
    for launch in (PAIR, FIXED8):
        with bqp_amd.Propagator(field, launch=launch, **kw) as p:
            y, s, info = p.propagate_batch_final(y0, t0, tf, bc)
        out.append((y, s, info))
    return out


def _assert_same(a, b):
    (ya, sa, ia), (yb, sb, ib) = a, b
    assert np.array_equal(sa, sb)
    assert np.array_equal(ya, yb, equal_nan=True)
    for k in ("steps_per_row", "evals_per_row"):
        assert np.array_equal(ia[k], ib[k]), k


@pytest.mark.parametrize("method,control,nmax,drag", [("rk87", "scaled", 40, "vallado"),
                                                      ("dopri5", "scaled", 20, "none"),
                                                      ("dop853", "scaled", 30, "vallado"),
                                                      ("dopri5", "compat", 12, "none")])
def test_pair_equals_fixed_g8(method, control, nmax, drag):
    n = 70_000  # more than one dealt round on the device: queue refills
    w = inputs.workload("w1", n=n, seed=41)
    rng = np.random.default_rng(41)
    tf = rng.uniform(600.0, 3 * 3600.0, n)
    announce(f"pair kernel vs fixed G=8: {method}/{control}, N={nmax}, drag={drag}, {n} rows, ragged 10 min-3 h")
    field = bqp_amd.kaula(nmax)
    bc = w["bc"] if drag != "none" else None
    a, b = _both(field, w["y0"], 0.0, tf, bc, method=method, control=control, rtol=1e-11, drag=drag)
    assert np.all(a[1] == 0)
    _assert_same(a, b)


def test_pair_encke_and_sampled_equal_fixed_g8():
    n = 3000
    w = inputs.workload("w1", n=n, seed=42)
    field = bqp_amd.kaula(30)
    announce("pair kernel vs fixed G=8: Encke RK8(7)13M 7 h with drag; sampled Cowell at 6 times")
    a, b = _both(field, w["y0"], 0.0, np.full(n, 7 * 3600.0), w["bc"], method="rk87", control="absolute",
                 eps=1e-6, drag="vallado", formulation="encke")
    assert np.all(a[1] == 0)
    _assert_same(a, b)
    te = np.array([0.0, 600.0, 1800.0, 3600.0, 5000.0, 9000.0])
    res = []
    for launch in (PAIR, FIXED8):
        with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, launch=launch) as p:
            res.append(p.propagate_batch_sampled(w["y0"][:500], 0.0, te))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("nmax", [40, 20])
def test_pair_split_equals_fixed_g8(nmax):
    """Split precision: degree 40 takes the rescaled FP32 table, degree 20
    the (a, b) table; both equal the fixed kernel at G = 8."""
    n = 70_000
    w = inputs.workload("w1", n=n, seed=45)
    rng = np.random.default_rng(45)
    tf = rng.uniform(600.0, 2 * 3600.0, n)
    announce(f"pair kernel vs fixed G=8: split precision, N={nmax}, drag, {n} rows, ragged 10 min-2 h")
    a, b = _both(bqp_amd.kaula(nmax), w["y0"], 0.0, tf, w["bc"], method="rk87", rtol=1e-11, drag="vallado",
                 precision="split")
    assert np.all(a[1] == 0)
    _assert_same(a, b)


def test_pair_failing_rows_and_tiny_batches():
    w = inputs.workload("w1", n=200, seed=43)
    field = bqp_amd.kaula(20)
    y0 = w["y0"].copy()
    y0[3, 2] = np.nan          # invalid input
    y0[70, :3] *= 0.5          # below the ground: a status, NaN-filled
    tf = np.full(200, 3600.0)
    tf[5] = 0.0                # zero span
    announce("pair kernel vs fixed G=8: invalid rows, zero span, max-steps, batches of 1, 63, 65 rows")
    a, b = _both(field, y0, 0.0, tf, method="rk87", rtol=1e-11, r_min=6378137.0, max_steps=300)
    _assert_same(a, b)
    assert a[1][3] != 0 and a[1][70] != 0 and a[1][5] == 0
    for n in (1, 63, 65):
        a, b = _both(field, w["y0"][:n], 0.0, np.full(n, 1800.0), method="rk87", rtol=1e-11)
        assert np.all(a[1] == 0)
        _assert_same(a, b)


def test_pair_kernel_vs_oracle(orc):
    n = 128
    w = inputs.workload("w1", n=n, seed=44)
    field = bqp_amd.kaula(40)
    announce("pair kernel vs oracle: 128 rows, N=40, Vallado drag, 6 h")
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado", launch=PAIR) as p:
        yg, sg, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 6 * 3600.0), w["bc"])
    _, model, integ = oracle_pair(orc, field, method="rk87", rtol=1e-12, drag="vallado")
    yo, so, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, 6 * 3600.0, w["bc"])
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    announce(f"  pair vs oracle max {err.max():.3e} m")
    assert err.max() < 1e-2


def test_launch_combinations_are_checked():
    field = bqp_amd.kaula(8)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        p.set_launch(**PAIR)
        for bad in ({"waves_per_group": 4, "teams_per_block": 2}, {"waves_per_group": 8, "teams_per_block": 4},
                    {"waves_per_group": 16, "teams_per_block": 1}):
            with pytest.raises(Exception, match="teams_per_block"):
                p.set_launch(**bad)


def test_automatic_launch_takes_the_pair_kernel_where_it_measured_faster():
    """The automatic launch (waves_per_group = teams_per_block = 0) takes the
    pair kernel for FP64, split and Encke fields of degree >= 60 whose batch
    is one round of G = 4 groups (65,536 rows at most on 256 CUs), the
    fixed-G kernel otherwise; the automatic pair call equals an explicit one
    (FP64) and the fixed kernel at G = 8 (split) bit for bit."""
    n = 65536
    w = inputs.workload("w1", n=n, seed=43)
    tf = np.full(n, 120.0)
    f70 = bqp_amd.kaula(70)
    announce("automatic launch: degree 70 / 20, 65,536 and 98,304 rows, FP64, Encke, split")
    with bqp_amd.Propagator(f70, method="rk87", rtol=1e-12, drag="vallado") as p:
        ya, sa, _ = p.propagate_batch_final(w["y0"], 0.0, tf, w["bc"])
        used = p.last_launch()
        assert used["kernel"] == "pair" and used["waves_per_group"] == 8 and used["teams"] == 2
        assert used["threads"] == 512 and used["blocks"] == n // 128
        p.set_launch(**PAIR)
        yb, sb, _ = p.propagate_batch_final(w["y0"], 0.0, tf, w["bc"])
        assert np.all(sa == 0) and np.array_equal(sa, sb) and np.array_equal(ya, yb)
        p.set_launch(waves_per_group=4)
        p.propagate_batch_final(w["y0"], 0.0, tf, w["bc"])
        assert p.last_launch()["kernel"] == "fixed" and p.last_launch()["waves_per_group"] == 4
    with bqp_amd.Propagator(f70, method="rk87", control="absolute", eps=1e-6, formulation="encke") as p:
        p.propagate_batch_final(w["y0"][:4096], 0.0, tf[:4096])
        assert p.last_launch()["kernel"] == "fixed"  # 64 groups: G = 16
        p.propagate_batch_final(w["y0"], 0.0, tf)
        assert p.last_launch()["kernel"] == "pair"
    with bqp_amd.Propagator(f70, method="rk87", rtol=1e-12, precision="split") as p:
        ya, sa, _ = p.propagate_batch_final(w["y0"], 0.0, tf)
        assert p.last_launch()["kernel"] == "pair"
        p.set_launch(waves_per_group=8)
        yb, sb, _ = p.propagate_batch_final(w["y0"], 0.0, tf)
        assert p.last_launch()["kernel"] == "fixed" and p.last_launch()["waves_per_group"] == 8
        assert np.all(sa == 0) and np.array_equal(sa, sb) and np.array_equal(ya, yb)
    with bqp_amd.Propagator(bqp_amd.kaula(20), method="rk87", rtol=1e-12) as p:
        p.propagate_batch_final(w["y0"], 0.0, tf)
        assert p.last_launch()["kernel"] == "fixed"
    w2 = inputs.workload("w1", n=98304, seed=44)
    with bqp_amd.Propagator(f70, method="rk87", rtol=1e-12) as p:
        p.propagate_batch_final(w2["y0"], 0.0, np.full(98304, 60.0))
        assert p.last_launch()["kernel"] == "fixed" and p.last_launch()["waves_per_group"] == 4


def test_automatic_pair_launch_under_graph_replay():
    """The automatic pair launch (degree 70, one round of 4-wave groups)
    captured in a HIP graph and replayed, with other device work and
    synchronisation between replays, gives the direct launch's bits."""
    from bqp_amd.device import DeviceArray, synchronize

    n = 40_000
    w = inputs.workload("w1", n=n, seed=46)
    announce("automatic pair launch under hipGraph replay: degree 70, 40,000 rows x 10 min, drag")
    y0 = DeviceArray.from_numpy(w["y0"])
    t0 = DeviceArray.from_numpy(np.zeros(n))
    tfd = DeviceArray.from_numpy(np.full(n, 600.0))
    bc = DeviceArray.from_numpy(w["bc"])
    pos = np.random.default_rng(1).normal(size=(4096, 3)) * 7e6
    with bqp_amd.Propagator(bqp_amd.kaula(70), method="rk87", rtol=1e-12, drag="vallado") as p:
        ref, sref, _ = p.propagate_batch_final_dev(y0, t0, tfd, bc)
        assert p.last_launch()["kernel"] == "pair"
        ref, sref = ref.numpy(), sref.numpy()
        p.set_launch(use_graph=1)
        yf = DeviceArray((n, 6))
        st = DeviceArray((n,), np.int32)
        for rep in range(3):
            yf.copy_from(np.zeros((n, 6)))
            p.gravity_accel(pos)  # other device work between replays
            synchronize()
            out, s, info = p.propagate_batch_final_dev(y0, t0, tfd, bc, yf, st)
            assert p.last_launch()["kernel"] == "pair"
            assert np.array_equal(out.numpy(), ref), rep
            assert np.array_equal(s.numpy(), sref) and info["rows_ok"] == n
