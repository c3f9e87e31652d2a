"""GPU propagation against the CPU oracle on the same seeded inputs.

Bars: final positions within 1 cm of the oracle at equal tolerances (the
required parity bar; observed differences are far smaller), bit-identical
results across launch shapes, NaN-filled failed rows with their status."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from gpu_util import announce, oracle_pair, pos_err

pytestmark = pytest.mark.gpu

DAY = 86400.0


def _run_both(orc, field, w, n, method, rtol, drag, tf=None, control="scaled", eps=1e-8, launch=None,
              h_max=0.0):
    y0 = w["y0"][:n]
    t0 = np.zeros(n)
    tf = w["tf"][:n] if tf is None else np.full(n, tf)
    bc = w["bc"][:n] if drag != "none" else None
    with bqp_amd.Propagator(field, method=method, rtol=rtol, drag=drag, control=control, eps=eps,
                            launch=launch, h_max=h_max) as p:
        yg, sg, info = p.propagate_batch_final(y0, t0, tf, bc)
    _, model, integ = oracle_pair(orc, field, method=method, rtol=rtol, drag=drag, control=control, eps=eps,
                                  h_max=h_max)
    yo, so, steps, evals = orc.propagate_batch(model, integ, y0, t0, tf, bc)
    return yg, sg, info, yo, so, steps


def test_w0_dopri5_vs_oracle(orc):
    """The CPU-runnable configuration in full: 1,024 seeded LEO orbits, one
    day, DOPRI5(4) at rtol = atol = 1e-10 (canonical units), degree/order 8
    Kaula field, no drag."""
    announce("w0 DOPRI5 N=8 1 day, 1,024 rows (the whole workload) vs oracle")
    w = inputs.workload("w0", n=1024)
    field = bqp_amd.kaula(8)
    yg, sg, info, yo, so, steps = _run_both(orc, field, w, 1024, "dopri5", 1e-10, "none")
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    assert err.max() < 1e-2, err.max()
    # steps agree with the restated controller (rare ties may flip a decision)
    assert np.mean(info["steps_per_row"] == steps) > 0.9


def _oracle_ulp_spread(orc, field, w, n, method, rtol, drag, tf, all_components=True, components=(0,)):
    """The oracle against itself with one component of the initial state moved
    by one ulp, each of the six components up and down (12 runs; with
    all_components=False only the listed components, up); per row the largest displacement: the
    conditioning floor any two correct implementations can be expected to
    meet.  One perturbation alone underestimates it on sensitive rows (a row
    whose x0 + 1 ulp moves it by 3 mm moves by 3 cm under vx + 1 ulp)."""
    _, model, integ = oracle_pair(orc, field, method=method, rtol=rtol, drag=drag)
    bc = w["bc"][:n] if drag != "none" else None
    y0 = w["y0"][:n]
    a, _, _, _ = orc.propagate_batch(model, integ, y0, 0.0, tf, bc)
    spread = np.zeros(n)
    for comp in (range(6) if all_components else components):
        for d in ((np.inf, -np.inf) if all_components else (np.inf,)):
            yp = y0.copy()
            yp[:, comp] = np.nextafter(yp[:, comp], d)
            b, _, _, _ = orc.propagate_batch(model, integ, yp, 0.0, tf, bc)
            spread = np.maximum(spread, pos_err(a, b))
    return spread


def test_rk87_drag_vs_oracle(orc):
    """With Vallado drag the piecewise scale heights make the final state
    sensitive at the centimetre level to ulp-sized differences (the oracle
    against itself with one-ulp changes of the initial state spreads by up to
    ~3 cm over a day), so the bar is: the median row agrees to 1e-5 m and the
    worst row stays within 3x the oracle's own one-ulp spread (+1 mm)."""
    announce("RK8(7)13M N=20 Vallado drag 1 day, 64 rows vs oracle")
    w = inputs.workload("w1", n=64)
    field = bqp_amd.kaula(20)
    yg, sg, info, yo, so, steps = _run_both(orc, field, w, 64, "rk87", 1e-12, "vallado", tf=DAY)
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    spread = _oracle_ulp_spread(orc, field, w, 64, "rk87", 1e-12, "vallado", DAY)
    announce(f"drag parity max {err.max():.3e} median {np.median(err):.3e}; oracle 1-ulp spread max {spread.max():.3e}")
    assert np.median(err) < 1e-5
    assert err.max() <= 3.0 * spread.max() + 1e-3


def test_dop853_vs_oracle(orc):
    announce("DOP853 N=12 6 h, 64 rows vs oracle")
    w = inputs.workload("w0", n=64)
    field = bqp_amd.kaula(12)
    yg, sg, info, yo, so, steps = _run_both(orc, field, w, 64, "dop853", 1e-11, "none", tf=6 * 3600.0)
    assert np.all(sg == 0)
    assert pos_err(yg, yo).max() < 1e-2


def test_compat_controller_vs_oracle(orc):
    announce("DOPRI5 compat controller (absolute eps) vs oracle")
    w = inputs.workload("w0", n=64)
    field = bqp_amd.kaula(5)
    yg, sg, info, yo, so, steps = _run_both(orc, field, w, 64, "dopri5", 1e-12, "none", tf=3 * 3600.0,
                                            control="compat", eps=1e-6, h_max=300.0)
    assert np.all(sg == 0)
    assert pos_err(yg, yo).max() < 1e-2


def test_w1_seven_day_n70_vs_oracle(orc):
    """The headline parity claim: 7-day LEO arcs at degree/order 70,
    RK8(7)13M rtol 1e-12, final positions within 1 cm of the oracle (gravity
    only; with drag the bar is the oracle's own 1-ulp spread, see above)."""
    announce("w1 RK8(7)13M N=70 7 days (gravity only), 16 rows vs oracle")
    w = inputs.workload("w1", n=16)
    field = bqp_amd.kaula(70)
    yg, sg, info, yo, so, steps = _run_both(orc, field, w, 16, "rk87", 1e-12, "none")
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    announce(f"w1 parity max |dr| = {err.max():.3e} m")
    assert err.max() < 1e-2


def test_w1_seven_day_n70_drag_vs_oracle_spread(orc):
    """The headline problem with drag: 64 rows strided across the 65,536-row
    workload (every altitude), 7 days, degree 70, Vallado drag, in the pair
    kernel (the automatic launch of the full batch; its numerics are G = 8's).
    The piecewise scale heights make the final state sensitive to ulp-level
    differences, so the yardstick is the oracle against itself with x0 or vy
    moved by one ulp: the worst row within 3x the worst spread plus 1 cm, and no
    more rows beyond 1 cm than twice the oracle's own count plus two."""
    announce("w1 N=70 drag 7 days, 64 strided rows (pair kernel) vs oracle and its 1-ulp spread")
    w = inputs.workload("w1", n=65536)
    idx = np.arange(0, 65536, 1024)
    n = idx.size
    ws = {"y0": w["y0"][idx], "bc": w["bc"][idx], "tf": np.full(n, 7 * DAY)}
    field = bqp_amd.kaula(70)
    yg, sg, info, yo, so, steps = _run_both(orc, field, ws, n, "rk87", 1e-12, "vallado",
                                            launch={"waves_per_group": 8, "teams_per_block": 2})
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    # x0 and vy (along-track for most rows) moved by one ulp, separately
    spread = _oracle_ulp_spread(orc, field, ws, n, "rk87", 1e-12, "vallado", 7 * DAY, all_components=False,
                                components=(0, 4))
    beyond, beyond_o = int(np.sum(err > 1e-2)), int(np.sum(spread > 1e-2))
    announce(f"w1 drag parity max {err.max():.3e} median {np.median(err):.3e}, {beyond}/{n} rows beyond 1 cm; "
             f"oracle 1-ulp spread (x0, vy) max {spread.max():.3e} median {np.median(spread):.3e}, "
             f"{beyond_o}/{n} beyond 1 cm")
    assert err.max() <= 3.0 * spread.max() + 1e-2
    assert beyond <= 2 * beyond_o + 2


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16])
def test_launch_shape_bit_identity(G):
    """For a given number of waves per trajectory group, results do not depend
    on grid size, hand-out order or queue order (bit-identical).  Different G
    reorder the gravity partial sums and reach the sectoral values by
    different products (1e-16-level differences in the acceleration); where
    such a difference flips a step-size decision the runs part by up to the
    local error tolerance, so across G rows agree to 10 rtol |r|."""
    announce(f"launch shape invariance G={G}")
    w = inputs.workload("w2", n=700)
    field = bqp_amd.kaula(10)
    tf = np.minimum(w["tf"][:700], 4 * 3600.0)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        p.set_launch(waves_per_group=G, sort_by_cost=0)
        ref, s_ref, _ = p.propagate_batch_final(w["y0"][:700], 0.0, tf)
        p.set_launch(waves_per_group=G, sort_by_cost=1, blocks=3)
        got, s_got, _ = p.propagate_batch_final(w["y0"][:700], 0.0, tf)
        p.set_launch(waves_per_group=1, sort_by_cost=1)
        g1, _, _ = p.propagate_batch_final(w["y0"][:700], 0.0, tf)
        # the learned cost model now orders the rows differently
        p.set_launch(waves_per_group=G, sort_by_cost=1)
        learned, s_learned, _ = p.propagate_batch_final(w["y0"][:700], 0.0, tf)
        p.set_launch(waves_per_group=G, sort_by_cost=2)
        by_rev, _, _ = p.propagate_batch_final(w["y0"][:700], 0.0, tf)
        learned_rows = p.cost_model()["rows"].sum()  # the three calls with sort_by_cost = 1
        assert learned_rows > 0 and learned_rows % 3 == 0
    assert np.array_equal(s_ref, s_got) and np.array_equal(s_ref, s_learned)
    assert np.array_equal(ref, got)
    assert np.array_equal(ref, learned)
    assert np.array_equal(ref, by_rev)
    r = np.linalg.norm(ref[:, :3], axis=1)
    assert np.all(pos_err(ref, g1) <= 10 * 1e-11 * r)


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16])
def test_fsal_every_g_matches_oracle(orc, G):
    """DOPRI5 reuses the last stage as the next step's first (FSAL): wave 0
    copies it into slot 0 of the group's LDS stage buffer only after every
    wave has read slot 0 for its solution sums.  Regression test for a race
    that let lagging waves read the copied value (four waves per SIMD at
    G = 16): every row must finish and match the oracle at every G."""
    announce(f"DOPRI5 compat FSAL, 16 rows x 3 h, N=8, G={G} vs oracle")
    field = bqp_amd.kaula(8)
    w = inputs.workload("w1", n=16, seed=8)
    span = 3 * 3600.0
    _, model, integ = oracle_pair(orc, field, method="dopri5", control="compat", eps=1e-9)
    yo, so, steps, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, span)
    with bqp_amd.Propagator(field, method="dopri5", control="compat", eps=1e-9) as p:
        p.set_launch(waves_per_group=G)
        yg, sg, info = p.propagate_batch_final(w["y0"], 0.0, span)
    assert np.all(sg == 0) and np.all(so == 0)
    assert pos_err(yg, yo).max() < 1e-2
    # same step sequences up to rare ulp-level accept/reject ties
    assert np.max(np.abs(info["steps_per_row"] - steps) / steps) <= 0.01


def test_batch_equals_single():
    announce("batch row == single-row call")
    w = inputs.workload("w0", n=130)
    field = bqp_amd.kaula(8)
    with bqp_amd.Propagator(field, method="dopri5", rtol=1e-10) as p:
        yb, _, _ = p.propagate_batch_final(w["y0"], 0.0, 3600.0)
        for i in (0, 64, 129):
            ys, _, _ = p.propagate_batch_final(w["y0"][i:i + 1], 0.0, 3600.0)
            assert np.array_equal(ys[0], yb[i])


def test_failed_rows_nan_filled_with_status():
    announce("row failure statuses")
    field = bqp_amd.kaula(8)
    w = inputs.workload("w0", n=8)
    y0 = w["y0"].copy()
    t0 = np.zeros(8)
    tf = np.full(8, 1800.0)
    y0[1, 2] = np.nan                 # invalid input
    tf[2] = -5.0                      # tf < t0: invalid input
    y0[3, :3] *= 0.5                  # starts below the minimum radius
    tf[4] = 0.0                       # zero span: returns y0 exactly
    with bqp_amd.Propagator(field, method="dopri5", rtol=1e-10, r_min=6378137.0) as p:
        yf, st, info = p.propagate_batch_final(y0, t0, tf)
    assert list(st) == [0, 5, 5, 6, 0, 0, 0, 0]
    assert np.all(np.isnan(yf[1])) and np.all(np.isnan(yf[2])) and np.all(np.isnan(yf[3]))
    assert np.array_equal(yf[4], y0[4])
    assert info["rows_failed"] == 3 and info["first_failed_row"] == 1
    with bqp_amd.Propagator(field, method="dopri5", rtol=1e-10, max_steps=5) as p:
        yf, st, _ = p.propagate_batch_final(w["y0"][:3], 0.0, 86400.0)
    assert np.all(st == 1) and np.all(np.isnan(yf))


def test_empty_batch():
    field = bqp_amd.kaula(4)
    with bqp_amd.Propagator(field) as p:
        yf, st, info = p.propagate_batch_final(np.zeros((0, 6)), np.zeros(0), np.zeros(0))
    assert yf.shape == (0, 6) and st.shape == (0,)


def test_sampled_matches_final():
    announce("sampled grid vs final-state path")
    w = inputs.workload("w0", n=32)
    field = bqp_amd.kaula(8)
    grid = np.array([0.0, 600.0, 1800.0, 5400.0])
    with bqp_amd.Propagator(field, method="dop853", rtol=1e-12) as p:
        out, st, _ = p.propagate_batch_sampled(w["y0"], 0.0, grid)
        assert np.all(st == 0)
        assert np.array_equal(out[:, 0], w["y0"])
        for k, t in enumerate(grid[1:], start=1):
            yf, _, _ = p.propagate_batch_final(w["y0"], 0.0, t)
            assert pos_err(out[:, k], yf).max() < 1e-4
        single, _, _ = p.propagate_batch_sampled(w["y0"], 0.0, grid[-1:])
        yf, _, _ = p.propagate_batch_final(w["y0"], 0.0, grid[-1])
        assert np.array_equal(single[:, 0], yf)


def test_split_precision_trajectory_error():
    """Split form vs FP64 GPU on 7-day N=70 arcs: final-position error <= 1 m."""
    announce("split vs FP64 N=70 7 days 64 rows")
    w = inputs.workload("w1", n=64)
    field = bqp_amd.kaula(70)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado") as p:
        y64, s64, _ = p.propagate_batch_final(w["y0"], 0.0, w["tf"][:64], w["bc"][:64])
        p.reconfigure(precision="split")
        y32, s32, _ = p.propagate_batch_final(w["y0"], 0.0, w["tf"][:64], w["bc"][:64])
    assert np.all(s64 == 0) and np.all(s32 == 0)
    err = pos_err(y32, y64)
    announce(f"split vs fp64 max |dr| = {err.max():.3e} m")
    assert err.max() < 1.0


def test_ragged_mixed_orbits_vs_oracle(orc):
    """Ragged batch: LEO/MEO/GEO rows with durations spanning >100x in one
    launch (work queue refills lanes), against the oracle."""
    announce("ragged w2 slice N=20 vs oracle")
    w = inputs.workload("w2", n=96, seed=3)
    field = bqp_amd.kaula(20)
    tf = np.minimum(w["tf"], 2 * DAY)  # bound the oracle's CPU time
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        yg, sg, info = p.propagate_batch_final(w["y0"], 0.0, tf)
    _, model, integ = oracle_pair(orc, field, method="rk87", rtol=1e-11)
    yo, so, steps, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, tf)
    assert np.all(sg == 0) and np.all(so == 0)
    assert steps.max() > 10 * max(1, steps.min())
    assert pos_err(yg, yo).max() < 1e-2


def test_graph_replay_matches_direct_launch():
    """A captured launch sequence replayed several times, with other device
    work and synchronisation between replays, gives the direct launch's bits."""
    announce("hipGraph replay vs direct")
    from bqp_amd.device import DeviceArray, synchronize

    n = 2000
    w = inputs.workload("w2", n=n, seed=5)
    field = bqp_amd.kaula(12)
    tf = np.minimum(w["tf"], 3 * 3600.0)
    y0 = DeviceArray.from_numpy(w["y0"])
    t0 = DeviceArray.from_numpy(np.zeros(n))
    tfd = DeviceArray.from_numpy(tf)
    pos = np.random.default_rng(0).normal(size=(4096, 3)) * 7e6
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        ref, sref, _ = p.propagate_batch_final_dev(y0, t0, tfd)
        ref, sref = ref.numpy(), sref.numpy()
        p.set_launch(use_graph=1)
        yf = DeviceArray((n, 6))
        st = DeviceArray((n,), np.int32)
        for rep in range(4):
            yf.copy_from(np.zeros((n, 6)))
            p.gravity_accel(pos)  # other device work between replays
            synchronize()
            out, s, info = p.propagate_batch_final_dev(y0, t0, tfd, None, yf, st)
            assert np.array_equal(out.numpy(), ref), rep
            assert np.array_equal(s.numpy(), sref)
            assert info["rows_ok"] == n


def test_graph_after_sort_buffer_regrowth_matches_direct_launch():
    """Regression: the captured sequence holds the hand-out tables' address.
    A graph call at 1,000 rows, then a direct call at 100,000 rows (which
    regrows and so frees that buffer), then the 1,000-row graph call again
    must recapture (or replay into live memory) and give the direct launch's
    bits, with the same per-row evaluation counts."""
    announce("hipGraph after the hand-out buffer regrows: graph n=1000, direct n=100000, graph n=1000")
    from bqp_amd.device import DeviceArray

    small, big = 1000, 100000
    w = inputs.workload("w1", n=big, seed=3)
    field = bqp_amd.kaula(8)
    tf_small = np.full(small, 1800.0)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        ref, sref, iref = p.propagate_batch_final(w["y0"][:small], 0.0, tf_small)
        y0 = DeviceArray.from_numpy(w["y0"][:small])
        t0 = DeviceArray.from_numpy(np.zeros(small))
        tfd = DeviceArray.from_numpy(tf_small)
        ev = DeviceArray((small,), np.int32)
        for rep in range(2):
            p.set_launch(use_graph=1)
            out, s, _ = p.propagate_batch_final_dev(y0, t0, tfd, None, None, None, None, ev)
            assert np.array_equal(out.numpy(), ref), rep
            assert np.array_equal(s.numpy(), sref)
            assert np.array_equal(ev.numpy(), iref["evals_per_row"])
            # a direct call on a larger batch regrows the hand-out buffer
            p.set_launch(use_graph=0)
            _, sb, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(big, 300.0))
            assert np.all(sb == 0)
            p.set_launch(use_graph=1)
            out, s, _ = p.propagate_batch_final_dev(y0, t0, tfd, None, None, None, None, ev)
            assert np.array_equal(out.numpy(), ref), rep
            assert np.array_equal(s.numpy(), sref)
            assert np.array_equal(ev.numpy(), iref["evals_per_row"])


def test_dealt_first_round_covers_every_row_once():
    """The first round of groups is dealt to the blocks by index (a
    back-and-forth deal over the hand-out order) and the queue starts after
    it.  At 70,000 rows and G = 4 the full grid deals four rounds over the
    CUs and queues the remaining 70 groups; grids of 700 and 3 blocks deal
    two rounds plus a partial one, and none.  Every row must run exactly once
    (status ok, finite state, its own evaluation count), and for one G the
    results must be bit-identical whatever the grid."""
    announce("dealt first round: 70,000 rows x 300 s, N=4, G=4, grids full / 700 / 3")
    w = inputs.workload("w1", n=70000)
    field = bqp_amd.kaula(4)
    out = {}
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-10) as p:
        for blocks in (0, 700, 3):
            p.set_launch(waves_per_group=4, blocks=blocks)
            y, st, info = p.propagate_batch_final(w["y0"], 0.0, np.full(70000, 300.0))
            out[blocks] = (y, st, info["evals_per_row"].copy())
    y0, s0, e0 = out[0]
    assert np.all(s0 == 0) and np.all(np.isfinite(y0)) and np.all(e0 > 0)
    for blocks in (700, 3):
        y, st, ev = out[blocks]
        assert np.array_equal(st, s0)
        assert np.array_equal(ev, e0)
        assert np.array_equal(y, y0)
