"""GPU tests of the ragged-batch workload at its full size and of the
launch-shape autotuner.

* The full 262,144-row ragged batch (60% LEO, 25% MEO, 15% GEO; durations
  log-uniform 1 h to 30 d; degree/order 20; RK8(7)13M rtol 1e-11) runs as one
  launch; 72 of its rows -- every 4,096th in duration order plus the eight
  longest 30-day rows, which run the whole launch -- are propagated by the
  oracle at their full durations.  Bar: 1 cm (gravity only); every row
  finishes; the picked rows' step counts span more than 100x.
* The autotuner's pilot measures every candidate shape, keeps the fastest,
  sets it, and caches it per workload signature; a propagation after tuning
  equals a propagation with that shape set by hand, bit for bit.
"""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from bqp_amd.autotune import autotune
from bqp_amd.device import DeviceArray
from gpu_util import announce, oracle_pair, pos_err

pytestmark = pytest.mark.gpu

DAY = 86400.0


def test_ragged_full_batch_full_durations_vs_oracle(orc):
    n = 262144
    w = inputs.workload("w2", n=n)
    tf = w["tf"]
    order = np.argsort(tf)
    pick = np.unique(np.concatenate([order[::4096], order[-8:]]))
    announce(f"ragged workload: full batch {n} rows, degree 20, rtol 1e-11; {pick.size} rows at full duration vs oracle")
    field = bqp_amd.kaula(20)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        p.set_launch(waves_per_group=8, sort_by_cost=1)
        y0 = DeviceArray.from_numpy(w["y0"])
        t0 = DeviceArray.from_numpy(np.zeros(n))
        tfd = DeviceArray.from_numpy(tf)
        yf, st, info = p.propagate_batch_final_dev(y0, t0, tfd)
        yg, sg = yf.numpy(), st.numpy()
    assert np.all(sg == 0) and info["rows_ok"] == n
    assert np.all(np.isfinite(yg))
    _, model, integ = oracle_pair(orc, field, method="rk87", rtol=1e-11)
    yo, so, steps, _ = orc.propagate_batch(model, integ, w["y0"][pick], 0.0, tf[pick])
    assert np.all(so == 0)
    assert tf[pick].max() > 29.9 * DAY and tf[pick].min() < 1.01 * 3600.0
    assert steps.max() > 100 * steps.min()
    err = pos_err(yg[pick], yo)
    announce(f"ragged workload vs oracle: max {err.max():.3e} m, median {np.median(err):.3e} m; "
             f"oracle steps {steps.min()}..{steps.max()}; kernel {info['kernel_ms']:.1f} ms")
    assert err.max() < 1e-2


def test_autotuner_picks_measured_fastest_and_caches(tmp_path):
    n = 8192
    w = inputs.workload("w2", n=n, seed=2)
    tf = np.minimum(w["tf"], 2 * DAY)
    field = bqp_amd.kaula(20)
    cache = str(tmp_path / "launch.json")
    announce(f"autotuner: pilot over {n} ragged rows, degree 20")
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as p:
        rep = autotune(p, w["y0"], 0.0, tf, cache_path=cache, use_cache=True)
        assert not rep["cached"]
        shapes = [t["launch"]["waves_per_group"] for t in rep["trials"]]
        assert sorted(set(shapes)) == [1, 2, 4, 8, 16]
        assert all(t["kernel_ms"] > 0.0 for t in rep["trials"])
        best = min(rep["trials"], key=lambda t: t["kernel_ms"])
        assert rep["launch"] == best["launch"]
        assert all(p.launch[k] == v for k, v in best["launch"].items())
        tuned, s1, _ = p.propagate_batch_final(w["y0"], 0.0, tf)
        rep2 = autotune(p, w["y0"], 0.0, tf, cache_path=cache, use_cache=True)
        assert rep2["cached"] and rep2["launch"] == rep["launch"]
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11) as q:
        q.set_launch(**rep["launch"])
        direct, s2, _ = q.propagate_batch_final(w["y0"], 0.0, tf)
    announce(f"autotuner chose {rep['launch']} ({rep['pilot_kernel_ms']:.2f} ms pilot)")
    assert np.all(s1 == 0) and np.array_equal(s1, s2)
    assert np.array_equal(tuned, direct)
