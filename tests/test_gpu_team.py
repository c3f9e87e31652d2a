"""GPU tests of the team kernel (bqp_team.hpp), taken by FP64 and Encke calls
that ask for it (launch teams_per_block 1, 2 or 4) on fields of degree >= 30.

* against the oracle (Cowell RK8(7)13M with and without drag, DOPRI5 with
  its FSAL reuse, sampled output, Encke) to the usual bars;
* its results do not depend on how the waves are divided: groups per
  workgroup 1, 2 and 4, grids from the full device down to a few blocks
  (queue refills), batch against single rows -- all bit-identical;
* the fixed-G kernel (teams_per_block = 0, the default) agrees to the
  tolerance level.
"""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from gpu_util import announce, oracle_pair, pos_err

pytestmark = pytest.mark.gpu

DAY = 86400.0
TEAM = {"teams_per_block": 4}


def test_team_kernel_vs_oracle_gravity_and_drag(orc):
    announce("team kernel: 256 rows, N=40, RK8(7)13M rtol 1e-12, 6 h, gravity only and Vallado drag, vs oracle")
    n = 256
    w = inputs.workload("w1", n=n, seed=31)
    field = bqp_amd.kaula(40)
    for drag in ("none", "vallado"):
        bc = w["bc"] if drag != "none" else None
        with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag=drag, launch=TEAM) as p:
            yg, sg, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 6 * 3600.0), bc)
            p.set_launch(teams_per_block=0)
            yfix, sfix, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 6 * 3600.0), bc)
        _, model, integ = oracle_pair(orc, field, method="rk87", rtol=1e-12, drag=drag)
        yo, so, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, 6 * 3600.0, bc)
        assert np.all(sg == 0) and np.all(so == 0) and np.all(sfix == 0)
        err = pos_err(yg, yo)
        announce(f"  drag={drag}: team vs oracle max {err.max():.3e} m, team vs fixed-G max "
                 f"{pos_err(yg, yfix).max():.3e} m")
        assert err.max() < 1e-2
        assert pos_err(yg, yfix).max() < 1e-2


def test_team_kernel_dopri5_fsal_vs_oracle(orc):
    announce("team kernel: DOPRI5 (FSAL), 128 rows, N=30, rtol 1e-10, 3 h vs oracle")
    n = 128
    w = inputs.workload("w0", n=n, seed=32)
    field = bqp_amd.kaula(30)
    with bqp_amd.Propagator(field, method="dopri5", rtol=1e-10, launch=TEAM) as p:
        yg, sg, info = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 3 * 3600.0))
    _, model, integ = oracle_pair(orc, field, method="dopri5", rtol=1e-10)
    yo, so, steps, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, 3 * 3600.0)
    assert np.all(sg == 0) and np.all(so == 0)
    assert pos_err(yg, yo).max() < 1e-2
    assert np.mean(info["steps_per_row"] == steps) > 0.9


def test_team_kernel_sampled_vs_oracle(orc):
    announce("team kernel: sampled output, 64 rows, N=36, 13 samples over 12 h vs oracle sampled")
    n = 64
    w = inputs.workload("w1", n=n, seed=33)
    field = bqp_amd.kaula(36)
    te = np.concatenate([[0.0, 0.5], np.arange(1, 12) * 3600.0 + 0.25, [12 * 3600.0]])
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, launch=TEAM) as p:
        yg, sg, _ = p.propagate_batch_sampled(w["y0"], 0.0, te)
    _, model, integ = oracle_pair(orc, field, method="rk87", rtol=1e-12)
    yo, so, _, _ = orc.propagate_sampled_batch(model, integ, w["y0"], 0.0, te)
    assert np.all(sg == 0) and np.all(so == 0)
    assert pos_err(yg, yo).max() < 1e-2


def test_team_kernel_encke_vs_oracle(orc):
    announce("team kernel: Encke (DOP853, absolute eps 1e-6), 128 rows, N=30, drag, 8 h vs oracle Encke")
    n = 128
    w = inputs.workload("w1", n=n, seed=34)
    field = bqp_amd.kaula(30)
    kw = dict(method="dop853", control="absolute", eps=1e-6, drag="vallado", formulation="encke", launch=TEAM)
    with bqp_amd.Propagator(field, **kw) as p:
        yg, sg, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 8 * 3600.0), w["bc"])
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    model = orc.make_model(f, drag="vallado", formulation="encke")
    integ = orc.make_integ(method="dop853", control="absolute", eps=1e-6)
    yo, so, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, 8 * 3600.0, w["bc"])
    assert np.all(sg == 0) and np.all(so == 0)
    err = pos_err(yg, yo)
    announce(f"  Encke team vs oracle max {err.max():.3e} m, median {np.median(err):.3e} m")
    assert np.median(err) < 1e-3 and err.max() < 5e-2


def test_team_results_do_not_depend_on_the_division():
    """70,000 rows (more than the 65,536 lanes of four groups on 256 CUs),
    300 s each, degree 30: one, two or four groups per workgroup and grids of
    the whole device, 100 blocks and 7 blocks (most rows then come from the
    queue) give the same bits, status and evaluation counts."""
    announce("team kernel division invariance: 70,000 rows x 300 s, N=30, teams 1/2/4 x grids full/100/7")
    n = 70000
    w = inputs.workload("w1", n=n, seed=35)
    field = bqp_amd.kaula(30)
    tf = np.full(n, 300.0)
    tf[::3] = 200.0
    ref = None
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11, drag="vallado") as p:
        for teams in (4, 2, 1):
            for blocks in (0, 100, 7):
                p.set_launch(teams_per_block=teams, blocks=blocks)
                y, st, info = p.propagate_batch_final(w["y0"], 0.0, tf, w["bc"])
                got = (y, st, info["evals_per_row"].copy(), info["steps_per_row"].copy())
                if ref is None:
                    ref = got
                    assert np.all(st == 0) and np.all(np.isfinite(y))
                    continue
                assert np.array_equal(got[1], ref[1]), (teams, blocks)
                assert np.array_equal(got[2], ref[2]), (teams, blocks)
                assert np.array_equal(got[3], ref[3]), (teams, blocks)
                assert np.array_equal(got[0], ref[0]), (teams, blocks)


def test_team_batch_rows_equal_single_rows():
    """A row propagated alone (one group, all 16 waves on it) equals the same
    row inside the 65,536-row batch (four groups per workgroup, division
    changing as groups finish), bit for bit; degree 70, Vallado drag."""
    announce("team kernel: single rows vs the same rows in a 65,536-row batch (N=70, drag, 1 h)")
    n = 65536
    w = inputs.workload("w1", n=n)
    field = bqp_amd.kaula(70)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado", launch=TEAM) as p:
        yb, sb, ib = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 3600.0), w["bc"])
        assert np.all(sb == 0)
        for i in (0, 1, 4097, 65535):
            y1, s1, i1 = p.propagate_batch_final(w["y0"][i:i + 1], 0.0, [3600.0], w["bc"][i:i + 1])
            assert s1[0] == 0
            assert np.array_equal(y1[0], yb[i]), i
            assert i1["evals_per_row"][0] == ib["evals_per_row"][i]
