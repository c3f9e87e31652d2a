"""Error behaviour of the C ABI through the Python mirror (a device context
is needed, so these are GPU tests): invalid configurations and call orders
fail with a message and leave the context usable."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from gpu_util import announce

pytestmark = pytest.mark.gpu


def test_invalid_configurations_rejected():
    announce("ABI errors: invalid configurations")
    f = bqp_amd.kaula(4)
    with pytest.raises(ValueError):
        bqp_amd.Propagator(f, formulation="kepler")
    with pytest.raises(ValueError):
        bqp_amd.Propagator(f, formulation="encke", precision="split")
    with pytest.raises(ValueError):
        bqp_amd.Propagator(f, rtol=-1.0)
    with bqp_amd.Propagator(f) as p:
        with pytest.raises(ValueError):
            p.reconfigure(formulation="encke", precision="surrogate")
        with pytest.raises(ValueError):
            p.set_launch(waves_per_group=3)
        # still usable afterwards
        w = inputs.workload("w0", n=8)
        yf, st, _ = p.propagate_batch_final(w["y0"], 0.0, 600.0)
        assert np.all(st == 0)


def test_surrogate_call_order_errors():
    announce("ABI errors: surrogate without set-up / packing")
    f = bqp_amd.kaula(8)
    w = inputs.workload("w0", n=8)
    with bqp_amd.Propagator(f, precision="surrogate") as p:
        with pytest.raises(ValueError, match="bqp_surrogate_init"):
            p.propagate_batch_final(w["y0"], 0.0, 600.0)
        with pytest.raises(ValueError):
            bqp_amd.Surrogate(p, nlow=9)  # nlow must be <= 8 and below the field degree
        sur = bqp_amd.Surrogate(p, nlow=4)
        with pytest.raises(ValueError, match="bqp_surrogate_pack"):
            p.propagate_batch_final(w["y0"], 0.0, 600.0)
        with pytest.raises(ValueError):
            sur.set_check(-1, 1e-9)
        with pytest.raises(ValueError):
            sur.set_params(np.zeros(10, np.float32))
        feat, targ, _ = sur.dataset(128, seed=1)
        with pytest.raises(ValueError):
            sur.infer(feat, "bf16")  # not packed yet
        with pytest.raises(ValueError):
            sur.pack("fp8")  # FP8 needs calibration rows
        sur.pack("bf16")
        sur.set_check(1, 0.0)
        yf, st, _ = p.propagate_batch_final(w["y0"], 0.0, 600.0)
        assert np.all(st == 0)


def test_encke_graph_replay_matches_direct_launch():
    announce("Encke under HIP-graph replay vs direct launches (bit-identical)")
    f = bqp_amd.kaula(8)
    w = inputs.workload("w1", n=256, seed=2)
    with bqp_amd.Propagator(f, method="rk87", control="absolute", eps=1e-7, formulation="encke") as p:
        a, sa, _ = p.propagate_batch_final(w["y0"], 0.0, 7200.0)
        p.set_launch(use_graph=1)
        b, sb, _ = p.propagate_batch_final(w["y0"], 0.0, 7200.0)
        c, sc, _ = p.propagate_batch_final(w["y0"], 0.0, 7200.0)  # replay
    assert np.all(sa == 0) and np.array_equal(a, b) and np.array_equal(b, c)
