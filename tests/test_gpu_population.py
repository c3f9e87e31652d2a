"""GPU test of the population driver: NSGA-II generations evaluated as
device batches give the objectives of the candidates' own propagations."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from bqp_amd.population import DepartureProblem, Nsga2
from gpu_util import announce

pytestmark = pytest.mark.gpu


def test_generations_are_device_batches():
    announce("population: 3 generations x 128 candidates, N=8, coast 0.05-0.2 d")
    w = inputs.workload("w1", n=128, seed=7)
    with bqp_amd.Propagator(bqp_amd.kaula(8), method="rk87", rtol=1e-10) as prop:
        prob = DepartureProblem(prop, w["y0"], None, 6378137.0 + 800e3, t_min_days=0.05, t_max_days=0.2)
        rng = np.random.Generator(np.random.PCG64(1))
        opt = Nsga2(prob, prob.lower, prob.upper, pop_size=128, seed=1, initial=prob.initial_population(128, rng))
        opt.run(2)
        assert opt.evaluations == 3 * 128 and prob.calls == 3
        assert np.all(np.isfinite(opt.F)) and np.all(opt.X >= prob.lower) and np.all(opt.X <= prob.upper)
        # the survivors' objectives are those of their own propagation
        y0, tf, bc = prob.candidates(opt.X)
        yf, st, _ = prop.propagate_batch_final(y0, 0.0, tf)
        assert np.all(st == 0)
        miss = np.abs(np.linalg.norm(yf[:, :3], axis=1) - prob.r_target) / 1e3
        assert np.allclose(opt.F[:, 1], miss, rtol=0, atol=1e-6)
        assert np.allclose(opt.F[:, 0], np.linalg.norm(opt.X[:, 1:4], axis=1))
