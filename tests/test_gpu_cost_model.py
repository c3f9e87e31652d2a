"""The learned cost model behind the longest-first hand-out order.

Each final-state call adds every row of at least one revolution that
finished ok to a table of (evaluations, revolutions, rows) over cells of
(altitude of the semi-major axis, cos i); the next call orders rows by revolutions x the learned
evaluations per revolution.  The order never changes results (bit identity
across orders is in test_gpu_propagate); these tests check that the table
holds exactly the sums the rows' own counts give, binned as documented in
include/bqp.h, and that it is forgotten when the configuration changes."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from bqp_amd import _native as nat

from gpu_util import announce

pytestmark = pytest.mark.gpu


def host_cells(y0, span, mu, radius):
    """Cell index and revolutions of every row (the binning of include/bqp.h)."""
    r = np.linalg.norm(y0[:, :3], axis=1)
    v2 = (y0[:, 3:] ** 2).sum(axis=1)
    en = 0.5 * v2 - mu / r
    a = -mu / (2.0 * en)
    rev = span * np.sqrt(mu / a**3)
    alt = np.maximum(a - radius, 1e5)
    ia = np.clip(np.floor(np.log10(alt / 1e5) * (nat.COST_ALT_BINS / 3.0)).astype(int), 0, nat.COST_ALT_BINS - 1)
    h = np.cross(y0[:, :3], y0[:, 3:])
    ci = h[:, 2] / np.linalg.norm(h, axis=1)
    ii = np.clip(np.floor(0.5 * (ci + 1.0) * nat.COST_INC_BINS).astype(int), 0, nat.COST_INC_BINS - 1)
    return ia, ii, rev


def test_cost_model_holds_the_rows_own_counts():
    announce("cost model: 3,000 mixed LEO/MEO/GEO rows x <= 2 h, N=20, sums vs row counts")
    w = inputs.workload("w2", n=3000)
    field = bqp_amd.kaula(20)
    tf = np.minimum(w["tf"], 2 * 3600.0)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        assert p.cost_model()["rows"].sum() == 0
        _, status, info = p.propagate_batch_final(w["y0"], 0.0, tf)
        cm = p.cost_model()
        ia, ii, rev = host_cells(w["y0"], tf, field.mu, field.radius)
        assert np.all(status == 0)
        ok = rev >= 1.0  # rows shorter than one revolution are not learned
        assert 0 < ok.sum() < len(ok)
        ev = np.zeros(cm["evals"].shape, dtype=np.int64)
        rv = np.zeros(cm["revolutions"].shape)
        rows = np.zeros(cm["rows"].shape, dtype=np.int64)
        np.add.at(ev, (ia[ok], ii[ok]), info["evals_per_row"][ok].astype(np.int64))
        np.add.at(rv, (ia[ok], ii[ok]), rev[ok])
        np.add.at(rows, (ia[ok], ii[ok]), 1)
        assert np.array_equal(cm["rows"], rows)
        assert np.array_equal(cm["evals"], ev)
        # revolutions are summed in units of 2^-16
        assert np.all(np.abs(cm["revolutions"] - rv) <= rows * 2.0**-16)
        # a second call adds to the sums; a sampled call and sort_by_cost 0 / 2 add nothing
        p.propagate_batch_final(w["y0"], 0.0, tf)
        assert np.array_equal(p.cost_model()["evals"], 2 * ev)
        p.propagate_batch_sampled(w["y0"][:100], 0.0, [600.0, 1200.0])
        p.set_launch(sort_by_cost=2)
        p.propagate_batch_final(w["y0"], 0.0, tf)
        p.set_launch(sort_by_cost=0)
        p.propagate_batch_final(w["y0"], 0.0, tf)
        assert np.array_equal(p.cost_model()["evals"], 2 * ev)
        # a new configuration forgets the table; setting the same one does not
        p.set_launch(sort_by_cost=1)
        p.propagate_batch_final(w["y0"], 0.0, tf)
        p.reconfigure(rtol=1e-11)
        assert p.cost_model()["rows"].sum() == 3 * ok.sum()
        p.reconfigure(rtol=1e-10)
        assert p.cost_model()["rows"].sum() == 0


def test_invalid_sort_mode_rejected():
    field = bqp_amd.kaula(4)
    with bqp_amd.Propagator(field) as p:
        with pytest.raises(Exception):
            p.set_launch(sort_by_cost=3)
