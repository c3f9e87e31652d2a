"""Host-side mirror of the reference's propagation interface over the HIP C ABI.

``Propagator`` owns one bqp context (one HIP stream, one device-resident
field) and exposes the batch call shapes:

* ``propagate_batch_final``   ragged rows (y0, t0, tf, Cd*A/m) -> final states,
  the shape of ``LightyearSession::integrate_variable_final_into``
  (dust_transfer/crates/lightyear_odeint_rs/src/session.rs:2325-2436);
* ``propagate_batch_sampled`` fixed sample grid -> (n, m, 6), the shape of
  ``integrate_batch_native_into`` (.../src/batch.rs:247-390) in inertial states;
* ``propagate_batch_final_dev`` the same on device-resident buffers;
* ``gravity_accel``          body-fixed accelerations
  (crates/satpy_core/src/gravity.rs:2638 ``spherical_gravity_impl_frame_packed``).

Errors follow the reference: bad arguments raise ``ValueError``; a device or
integration failure raises ``RuntimeError`` (``BqpError``); batch rows that fail
are NaN-filled and reported in the returned status array.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat
from .elements import MU_EARTH, OMEGA_EARTH, R_EARTH
from .field import Field
from .tableau_struct import Tableau, to_dict


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check_dev(buf, name, n, tail, dtype, optional=False):
    """Shape and dtype contract of a device buffer passed to a device-pointer
    entry point: the kernel indexes it as (n,) + tail elements of `dtype`
    with no bounds of its own, so a mismatch must be refused on the host.
    Returns the buffer's row count.  Raises ValueError."""
    if buf is None:
        if optional:
            return n
        raise ValueError(f"{name} is required")
    shape = getattr(buf, "shape", None)
    dt = getattr(buf, "dtype", None)
    if shape is None or dt is None or not hasattr(buf, "data_ptr"):
        raise ValueError(f"{name} must be a device buffer with shape, dtype and data_ptr()")
    shape = tuple(int(s) for s in shape)
    if len(shape) != 1 + len(tail) or shape[1:] != tuple(tail) or (n is not None and shape[0] != n):
        want = ("n" if n is None else str(n),) + tuple(str(t) for t in tail)
        raise ValueError(f"{name} must have shape ({', '.join(want)}), got {shape}")
    if np.dtype(dt) != np.dtype(dtype):
        raise ValueError(f"{name} must be {np.dtype(dtype).name}, got {np.dtype(dt).name}")
    return shape[0]


def tableau(method: str) -> dict:
    t = Tableau()
    rc = nat.load().bqp_get_tableau(nat.METHODS[method], C.byref(t))
    if rc != 0:
        raise ValueError(method)
    return to_dict(t)


class Propagator:
    def __init__(self, field: Field, *, method: str = "rk87", control: str = "scaled",
                 precision: str = "fp64", drag: str = "none", rtol: float = 1e-12,
                 atol_pos: float | None = None, atol_vel: float | None = None, eps: float = 1e-8,
                 h0: float = 0.0, h_min: float = 0.0, h_max: float = 0.0, max_steps: int = 500000,
                 max_rejects: int = 50, omega: float = OMEGA_EARTH, theta0: float = 0.0,
                 r_atm: float = R_EARTH, r_min: float = 0.0, launch: dict | None = None,
                 formulation: str = "cowell", ground_flattening: float = 0.0):
        self._lib = nat.load()
        self._ctx = C.c_void_p()
        self.field = field
        cfg = self._make_config(method, control, precision, drag, rtol, atol_pos, atol_vel, eps, h0, h_min,
                                h_max, max_steps, max_rejects, omega, theta0, r_atm, r_min, field, formulation,
                                ground_flattening)
        rc = self._lib.bqp_create(C.byref(cfg), C.byref(self._ctx))
        if rc != 0:
            if rc == nat.E_INVALID:
                raise ValueError("invalid propagator configuration")
            raise nat.BqpError(f"bqp_create failed ({rc}): no usable HIP device?")
        self.config = cfg
        self._check(self._lib.bqp_load_field(self._ctx, field.nmax, field.mmax, _dp(field.cbar), _dp(field.sbar),
                                             field.mu, field.radius))
        if launch:
            self.set_launch(**launch)

    # ------------------------------------------------------------ plumbing
    @staticmethod
    def _make_config(method, control, precision, drag, rtol, atol_pos, atol_vel, eps, h0, h_min, h_max,
                     max_steps, max_rejects, omega, theta0, r_atm, r_min, field, formulation="cowell",
                     ground_flattening=0.0) -> nat.Config:
        for name, table, val in (("method", nat.METHODS, method), ("control", nat.CONTROLS, control),
                                 ("precision", nat.PRECISIONS, precision), ("drag", nat.DRAGS, drag),
                                 ("formulation", nat.FORMULATIONS, formulation)):
            if val not in table:
                raise ValueError(f"unknown {name} {val!r}; use one of {sorted(table)}")
        du = field.radius if field is not None else R_EARTH
        mu = field.mu if field is not None else MU_EARTH
        vu = float(np.sqrt(mu / du))
        cfg = nat.Config()
        cfg.method = nat.METHODS[method]
        cfg.control = nat.CONTROLS[control]
        cfg.precision = nat.PRECISIONS[precision]
        cfg.drag = nat.DRAGS[drag]
        cfg.formulation = nat.FORMULATIONS[formulation]
        cfg.rtol = rtol
        cfg.atol_pos = rtol * du if atol_pos is None else atol_pos
        cfg.atol_vel = rtol * vu if atol_vel is None else atol_vel
        cfg.eps = eps
        cfg.h0, cfg.h_min, cfg.h_max = h0, h_min, h_max
        cfg.max_steps, cfg.max_rejects = int(max_steps), int(max_rejects)
        cfg.omega, cfg.theta0, cfg.r_atm, cfg.r_min = omega, theta0, r_atm, r_min
        cfg.ground_flattening = ground_flattening
        return cfg

    def _check(self, rc: int, allow_rows: bool = False):
        if rc == 0 or (allow_rows and rc == nat.E_ROWS_FAILED):
            return
        msg = self._lib.bqp_last_error(self._ctx).decode()
        if rc == nat.E_INVALID:
            raise ValueError(msg)
        raise nat.BqpError(f"bqp error {rc}: {msg}")

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.bqp_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_launch(self, waves_per_group: int = 0, use_graph: int = 0, blocks: int = 0, sort_by_cost: int = 1,
                   teams_per_block: int = 0):
        """Launch shape: waves per trajectory group G (0 = auto), HIP-graph
        replay, persistent grid size (0 = fill the device), hand-out order
        (0 input order, 1 longest first by the learned cost model, 2 longest
        first by revolutions), and the groups per workgroup (0 = the
        fixed-G kernel; 1, 2 or 4 with G = 0 = the team kernel, whose
        results do not depend on the launch shape or the batch, FP64 or
        Encke, fields of degree >= 30; 2 with G = 8 = the pair kernel, two
        groups per 8-wave workgroup whose results equal G = 8's, FP64, split
        or Encke).  The automatic launch (G = 0, groups 0) takes the pair
        kernel for fields of degree >= 60 whose batch is one round of
        4-wave groups; `last_launch` reports what a call ran."""
        L = nat.Launch(waves_per_group, use_graph, blocks, sort_by_cost, teams_per_block)
        self._check(self._lib.bqp_set_launch(self._ctx, C.byref(L)))
        self.launch = {"waves_per_group": waves_per_group, "use_graph": use_graph, "blocks": blocks,
                       "sort_by_cost": sort_by_cost, "teams_per_block": teams_per_block}

    def last_launch(self) -> dict:
        """Shape of the last propagation call: the persistent kernel that ran
        ("fixed", "team" or "pair"; "none" before any call), its group width,
        groups per workgroup, grid and workgroup size."""
        info = nat.LaunchInfo()
        self._check(self._lib.bqp_last_launch(self._ctx, C.byref(info)))
        return {"kernel": nat.KERNELS.get(info.kernel, str(info.kernel)), "waves_per_group": info.waves_per_group,
                "teams": info.teams, "blocks": info.blocks, "threads": info.threads}

    def cost_model(self) -> dict:
        """The learned cost model (sort_by_cost = 1): per (altitude bin,
        cos i bin) cell, the evaluations, revolutions and rows summed over
        the rows that finished ok since the field or configuration was set."""
        cells = nat.COST_ALT_BINS * nat.COST_INC_BINS
        ev = np.zeros(cells, dtype=np.int64)
        rev = np.zeros(cells)
        rows = np.zeros(cells, dtype=np.int64)
        self._check(self._lib.bqp_get_cost_model(self._ctx, ev.ctypes.data_as(C.POINTER(C.c_int64)),
                                                 rev.ctypes.data_as(C.POINTER(C.c_double)),
                                                 rows.ctypes.data_as(C.POINTER(C.c_int64))))
        shape = (nat.COST_ALT_BINS, nat.COST_INC_BINS)
        return {"evals": ev.reshape(shape), "revolutions": rev.reshape(shape), "rows": rows.reshape(shape)}

    def reconfigure(self, **kw):
        """Change integrator/model settings (same keywords as the constructor)."""
        cur = {k: getattr(self.config, k) for k, _ in nat.Config._fields_}
        names = {"method": nat.METHODS, "control": nat.CONTROLS, "precision": nat.PRECISIONS, "drag": nat.DRAGS,
                 "formulation": nat.FORMULATIONS}
        for k, v in kw.items():
            if k in names:
                if v not in names[k]:
                    raise ValueError(f"unknown {k} {v!r}")
                cur[k] = names[k][v]
            elif k in cur:
                cur[k] = v
            else:
                raise ValueError(f"unknown setting {k!r}")
        cfg = nat.Config(**cur)
        self._check(self._lib.bqp_set_config(self._ctx, C.byref(cfg)))
        self.config = cfg

    @property
    def flops_per_rhs(self) -> float:
        return float(self._lib.bqp_flops_per_rhs(self._ctx))

    # ------------------------------------------------------------ calls
    def gravity_accel(self, pos_bf) -> np.ndarray:
        pos = np.ascontiguousarray(pos_bf, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(pos)
        self._check(self._lib.bqp_gravity_accel_batch(self._ctx, _dp(pos), pos.shape[0], _dp(out)))
        return out

    def propagate_batch_final(self, y0, t0, tf, bc=None):
        """Rows (n,6) SI from t0[i] to tf[i]; returns (yf, status, info)."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
        n = y0.shape[0]
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
        tf = np.ascontiguousarray(np.broadcast_to(np.asarray(tf, dtype=np.float64), (n,)))
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(np.broadcast_to(np.asarray(bc, dtype=np.float64), (n,)))
            bcp = _dp(bc)
        yf = np.empty_like(y0)
        status = np.empty(n, dtype=np.int32)
        steps = np.empty(n, dtype=np.int32)
        evals = np.empty(n, dtype=np.int32)
        st = nat.Stats()
        rc = self._lib.bqp_propagate_batch_final(self._ctx, _dp(y0), _dp(t0), _dp(tf), bcp, n, _dp(yf),
                                                 _ip(status), _ip(steps), _ip(evals), C.byref(st))
        self._check(rc, allow_rows=True)
        info = st.as_dict()
        info["steps_per_row"] = steps
        info["evals_per_row"] = evals
        return yf, status, info

    def propagate_batch_final_dev(self, y0, t0, tf, bc=None, yf=None, status=None, steps=None, evals=None):
        """Device-resident variant: arguments are device buffers exposing
        ``data_ptr()`` (``bqp_amd.device.DeviceArray``; float64 states and
        times, int32 counters).  No host staging.  Returns (yf, status, stats)."""
        from .device import DeviceArray

        n = _check_dev(y0, "y0", None, (6,), np.float64)
        _check_dev(t0, "t0", n, (), np.float64)
        _check_dev(tf, "tf", n, (), np.float64)
        _check_dev(bc, "bc", n, (), np.float64, optional=True)
        if yf is None:
            yf = DeviceArray((n, 6), np.float64)
        if status is None:
            status = DeviceArray((n,), np.int32)
        _check_dev(yf, "yf", n, (6,), np.float64)
        _check_dev(status, "status", n, (), np.int32)
        _check_dev(steps, "steps", n, (), np.int32, optional=True)
        _check_dev(evals, "evals", n, (), np.int32, optional=True)
        st = nat.Stats()

        def ptr(x):
            return C.c_void_p(x.data_ptr()) if x is not None else None

        rc = self._lib.bqp_propagate_batch_final_dev(self._ctx, ptr(y0), ptr(t0), ptr(tf), ptr(bc), n, ptr(yf),
                                                     ptr(status), ptr(steps), ptr(evals), C.byref(st))
        self._check(rc, allow_rows=True)
        return yf, status, st.as_dict()

    def propagate_batch_sampled(self, y0, t0: float, t_eval, bc=None):
        """Rows (n,6) from a common t0, sampled at t_eval (m,); returns
        (states (n, m, 6), status, info)."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
        n = y0.shape[0]
        te = np.ascontiguousarray(t_eval, dtype=np.float64).reshape(-1)
        m = te.shape[0]
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(np.broadcast_to(np.asarray(bc, dtype=np.float64), (n,)))
            bcp = _dp(bc)
        out = np.empty((n, m, 6))
        status = np.empty(n, dtype=np.int32)
        st = nat.Stats()
        rc = self._lib.bqp_propagate_batch_sampled(self._ctx, _dp(y0), float(t0), _dp(te), m, bcp, n, _dp(out),
                                                   _ip(status), C.byref(st))
        self._check(rc, allow_rows=True)
        return out, status, st.as_dict()
