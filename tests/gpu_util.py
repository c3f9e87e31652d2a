"""Helpers shared by the GPU tests: build a product Propagator and the
matching oracle model for the same workload."""
import numpy as np

import bqp_amd
from bqp_amd import inputs
from bqp_amd.elements import MU_EARTH, R_EARTH

GM = MU_EARTH
RE = R_EARTH


def announce(msg: str):
    print(f"[gpu-test] {msg}", flush=True)


def product(field, **kw):
    return bqp_amd.Propagator(field, **kw)


def oracle_pair(orc, field, method="rk87", control="scaled", drag="none", rtol=1e-12, eps=1e-8, h_max=0.0,
                r_min=0.0, max_steps=500000, ground_flattening=0.0):
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius, mmax=field.mmax)
    model = orc.make_model(f, drag=drag, r_min=r_min, ground_flattening=ground_flattening)
    integ = orc.make_integ(method=method, control=control, rtol=rtol, eps=eps, h_max=h_max, max_steps=max_steps,
                           canonical=(field.radius, field.mu))
    return f, model, integ


def workload_slice(name, n, seed=0):
    w = inputs.workload(name, n=n, seed=seed)
    return w


def pos_err(a, b):
    return np.linalg.norm(a[..., :3] - b[..., :3], axis=-1)
