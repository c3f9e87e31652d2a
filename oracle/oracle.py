"""ctypes binding of the CPU FP64 oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg load this as the checker / CPU baseline.  The product
package ``bqp_amd`` never imports it.

The C sources restate the reference algorithm; see ``bqp_oracle.h`` for the
per-function citations and pinning status.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

from bqp_amd import tableau_struct as _ts

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK, MAX_STEPS, STEP_UNDERFLOW, NON_FINITE, MAX_REJECTS, INVALID_INPUT, BELOW_MIN_RADIUS = range(7)
CONTROL = {"scaled": 0, "absolute": 1, "compat": 2}
DRAG = {"none": 0, "vallado": 1, "compat": 2}
GRAVITY = {"normalised": 0, "compat": 1}
METHOD = {"dopri5": 0, "dop853": 1, "rk87": 2}

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Model(C.Structure):
    _fields_ = [
        ("field", C.c_void_p),
        ("omega", C.c_double),
        ("theta0", C.c_double),
        ("drag", C.c_int),
        ("r_atm", C.c_double),
        ("r_min", C.c_double),
        ("formulation", C.c_int),
        ("ground_flattening", C.c_double),
    ]


class Integ(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("control", C.c_int),
        ("rtol", C.c_double),
        ("atol_pos", C.c_double),
        ("atol_vel", C.c_double),
        ("eps", C.c_double),
        ("h0", C.c_double),
        ("h_min", C.c_double),
        ("h_max", C.c_double),
        ("max_steps", C.c_int64),
        ("max_rejects", C.c_int32),
        ("no_hcarry", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("rejects", C.c_int64),
        ("evals", C.c_int64),
        ("status", C.c_int32),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = bind(_LIB_PATH)
    return _lib


@contextlib.contextmanager
def variant(path: str):
    """Runs the wrapper's calls against another build of the oracle source
    (a copy with a mistake planted in it, built by a test) for the duration
    of the block; Field objects made inside it belong to that build."""
    global _lib
    saved = lib()
    _lib = bind(path)
    try:
        yield _lib
    finally:
        _lib = saved


def bind(path: str):
    """Loads an oracle library and declares its C signatures."""
    L = C.CDLL(path)
    L.orc_field_create.restype = C.c_void_p
    L.orc_field_create.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int]
    L.orc_field_destroy.argtypes = [C.c_void_p]
    L.orc_gravity_bf.argtypes = [C.c_void_p, _dp, _dp]
    L.orc_gravity_bf_batch.argtypes = [C.c_void_p, _dp, C.c_int64, _dp, C.c_int]
    L.orc_density.restype = C.c_double
    L.orc_density.argtypes = [C.c_int, C.c_double]
    L.orc_rhs.argtypes = [C.POINTER(Model), C.c_double, C.c_double, _dp, _dp]
    L.orc_propagate.argtypes = [C.POINTER(Model), C.POINTER(Integ), C.c_double, _dp,
                                C.c_double, C.c_double, _dp, C.POINTER(Stats)]
    L.orc_propagate_batch.argtypes = [C.POINTER(Model), C.POINTER(Integ), _dp, _dp, _dp, _dp,
                                      C.c_int64, _dp, _i32p, _i64p, _i64p, C.c_int]
    L.orc_kep2eci.argtypes = [C.c_double, _dp, _dp]
    L.orc_tableau.argtypes = [C.c_int, C.c_void_p]
    L.orc_two_body.argtypes = [C.c_double, _dp, C.c_double, _dp, _dp]
    L.orc_battin.argtypes = [C.c_double, _dp, _dp, _dp]
    L.orc_trace_begin.argtypes = [_dp, C.c_int64]
    L.orc_trace_end.restype = C.c_int64
    L.orc_propagate_sampled_batch.argtypes = [C.POINTER(Model), C.POINTER(Integ), _dp, C.c_double, _dp,
                                              C.c_int64, _dp, C.c_int64, _dp, _i32p, _i64p, _i64p, C.c_int]
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


class Field:
    """A prepared field (owns the C object)."""

    def __init__(self, cbar, sbar, nmax, mu, radius, mmax=None, form="normalised"):
        self.cbar = np.ascontiguousarray(cbar, dtype=np.float64)
        self.sbar = np.ascontiguousarray(sbar, dtype=np.float64)
        self.nmax = int(nmax)
        self.mmax = int(nmax if mmax is None else mmax)
        self.mu, self.radius = float(mu), float(radius)
        need = (self.nmax + 1) * (self.nmax + 2) // 2
        if self.cbar.size < need or self.sbar.size < need:
            raise ValueError("coefficient arrays shorter than the triangular size")
        self.ptr = lib().orc_field_create(self.nmax, self.mmax, _p(self.cbar), _p(self.sbar),
                                          self.mu, self.radius, GRAVITY[form])
        if not self.ptr:
            raise ValueError("orc_field_create rejected the field")

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_field_destroy(self.ptr)
            self.ptr = None

    def accel(self, pos, threads=0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(pos)
        lib().orc_gravity_bf_batch(self.ptr, _p(pos), pos.shape[0], _p(out), int(threads))
        return out


def density(model: str, h_m: float) -> float:
    return lib().orc_density(DRAG[model], float(h_m))


FORMULATION = {"cowell": 0, "encke": 1, "test_exp": 2}


def make_model(field: Field, omega=7.292115e-5, theta0=0.0, drag="none", r_atm=6378137.0,
               r_min=0.0, formulation="cowell", ground_flattening=0.0) -> Model:
    m = Model(field.ptr, omega, theta0, DRAG[drag], r_atm, r_min, FORMULATION[formulation], ground_flattening)
    m._field_ref = field  # the C field must outlive every model that points at it
    return m


def make_integ(method="rk87", control="scaled", rtol=1e-12, atol_pos=None, atol_vel=None,
               eps=1e-8, h0=0.0, h_min=0.0, h_max=0.0, max_steps=500000, max_rejects=50,
               canonical=(6378137.0, 3.986004418e14), hcarry=True) -> Integ:
    """Tolerances; atol defaults to rtol in canonical units (DU = R, TU = sqrt(R^3/mu))."""
    du, mu = canonical
    vu = np.sqrt(mu / du)
    if atol_pos is None:
        atol_pos = rtol * du
    if atol_vel is None:
        atol_vel = rtol * vu
    return Integ(METHOD[method], CONTROL[control], rtol, atol_pos, atol_vel, eps, h0, h_min,
                 h_max, max_steps, max_rejects, 0 if hcarry else 1)


def rhs(model: Model, bc, t, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    dy = np.empty(6)
    lib().orc_rhs(C.byref(model), float(bc), float(t), _p(y), _p(dy))
    return dy


def propagate(model: Model, integ: Integ, y0, t0, tf, bc=0.0):
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    yf = np.empty(6)
    st = Stats()
    lib().orc_propagate(C.byref(model), C.byref(integ), float(bc), _p(y0), float(t0), float(tf),
                        _p(yf), C.byref(st))
    return yf, st


def propagate_traced(model: Model, integ: Integ, y0, t0, tf, bc=0.0, cap=200000):
    """orc_propagate with the per-attempt trace on: returns (yf, stats,
    trace) with trace an (attempts, 6) array of t, h, err_norm, threshold,
    accepted, next h."""
    buf = np.zeros((cap, 6))
    lib().orc_trace_begin(_p(buf), cap)
    try:
        yf, st = propagate(model, integ, y0, t0, tf, bc)
    finally:
        n = lib().orc_trace_end()
    if n > cap:
        raise ValueError(f"trace overflow: {n} attempts > cap {cap}")
    return yf, st, buf[:n].copy()


def propagate_batch(model: Model, integ: Integ, y0, t0, tf, bc=None, threads=0):
    y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
    n = y0.shape[0]
    t0 = np.ascontiguousarray(np.broadcast_to(t0, (n,)), dtype=np.float64)
    tf = np.ascontiguousarray(np.broadcast_to(tf, (n,)), dtype=np.float64)
    bcp = None
    if bc is not None:
        bc = np.ascontiguousarray(np.broadcast_to(bc, (n,)), dtype=np.float64)
        bcp = _p(bc)
    yf = np.empty_like(y0)
    status = np.empty(n, dtype=np.int32)
    steps = np.empty(n, dtype=np.int64)
    evals = np.empty(n, dtype=np.int64)
    lib().orc_propagate_batch(C.byref(model), C.byref(integ), _p(y0), _p(t0), _p(tf), bcp, n,
                              _p(yf), status.ctypes.data_as(_i32p), steps.ctypes.data_as(_i64p),
                              evals.ctypes.data_as(_i64p), int(threads))
    return yf, status, steps, evals


def propagate_sampled_batch(model: Model, integ: Integ, y0, t0, t_eval, bc=None, threads=0):
    """Rows from a common t0 sampled at the increasing times t_eval (steps
    shortened onto each sample, controller carried through them); returns
    (states (n, m, 6), status, steps, evals)."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
    n = y0.shape[0]
    te = np.ascontiguousarray(t_eval, dtype=np.float64).reshape(-1)
    m = te.shape[0]
    bcp = None
    if bc is not None:
        bc = np.ascontiguousarray(np.broadcast_to(bc, (n,)), dtype=np.float64)
        bcp = _p(bc)
    out = np.empty((n, m, 6))
    status = np.empty(n, dtype=np.int32)
    steps = np.empty(n, dtype=np.int64)
    evals = np.empty(n, dtype=np.int64)
    lib().orc_propagate_sampled_batch(C.byref(model), C.byref(integ), _p(y0), float(t0), _p(te), m, bcp, n,
                                      _p(out), status.ctypes.data_as(_i32p), steps.ctypes.data_as(_i64p),
                                      evals.ctypes.data_as(_i64p), int(threads))
    return out, status, steps, evals


def kep2eci(mu, kep):
    kep = np.ascontiguousarray(kep, dtype=np.float64).reshape(-1, 6)
    out = np.empty_like(kep)
    for i in range(kep.shape[0]):
        lib().orc_kep2eci(float(mu), _p(kep[i]), _p(out[i]))
    return out


def two_body(mu, s0, dt, x=None):
    """Two-body state dt after s0; returns (state, eccentric-anomaly change)."""
    s0 = np.ascontiguousarray(s0, dtype=np.float64)
    out = np.empty(6)
    xx = np.array([np.nan if x is None else float(x)])
    rc = lib().orc_two_body(float(mu), _p(s0), float(dt), _p(out), _p(xx))
    if rc != 0:
        raise ValueError("baseline is not an ellipse or the Kepler solve did not converge")
    return out, float(xx[0])


def battin(mu, rb, dr):
    rb = np.ascontiguousarray(rb, dtype=np.float64)
    dr = np.ascontiguousarray(dr, dtype=np.float64)
    acc = np.zeros(3)
    lib().orc_battin(float(mu), _p(rb), _p(dr), _p(acc))
    return acc


def tableau(method: str) -> dict:
    t = _ts.Tableau()
    rc = lib().orc_tableau(METHOD[method], C.byref(t))
    if rc != 0:
        raise ValueError(method)
    return _ts.to_dict(t)
