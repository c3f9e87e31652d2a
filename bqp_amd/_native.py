"""ctypes binding of ``libbqp.so`` (the C ABI declared in include/bqp.h).

There is no CPU fallback: if the HIP library is missing or no GPU is present,
every call raises.  ``build()`` (bqp_amd.build) compiles the library in-tree.
"""
from __future__ import annotations

import ctypes as C
import os

from .tableau_struct import Tableau

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbqp.so")

OK, E_INVALID, E_NO_FIELD, E_DEVICE, E_ROWS_FAILED, E_UNSUPPORTED = range(6)
ROW_STATUS = {
    0: "ok",
    1: "max_steps",
    2: "step_underflow",
    3: "non_finite",
    4: "max_rejects",
    5: "invalid_input",
    6: "below_min_radius",
}
# cells of the learned cost model (include/bqp.h BQP_COST_ALT_BINS, BQP_COST_INC_BINS)
COST_ALT_BINS = 96
COST_INC_BINS = 16

METHODS = {"dopri5": 0, "dop853": 1, "rk87": 2, "dopri87": 2, "rk8(7)13m": 2}
CONTROLS = {"scaled": 0, "absolute": 1, "compat": 2}
DRAGS = {"none": 0, "vallado": 1, "compat": 2}
PRECISIONS = {"fp64": 0, "split": 1, "surrogate": 2}
FORMULATIONS = {"cowell": 0, "encke": 1}
MLP_FORMATS = {"fp32": 0, "bf16": 1, "fp8": 2}


class Config(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("control", C.c_int32),
        ("precision", C.c_int32),
        ("drag", C.c_int32),
        ("rtol", C.c_double),
        ("atol_pos", C.c_double),
        ("atol_vel", C.c_double),
        ("eps", C.c_double),
        ("h0", C.c_double),
        ("h_min", C.c_double),
        ("h_max", C.c_double),
        ("max_steps", C.c_int64),
        ("max_rejects", C.c_int32),
        ("formulation", C.c_int32),
        ("omega", C.c_double),
        ("theta0", C.c_double),
        ("r_atm", C.c_double),
        ("r_min", C.c_double),
        ("ground_flattening", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("rejects", C.c_int64),
        ("evals", C.c_int64),
        ("rows_ok", C.c_int64),
        ("rows_failed", C.c_int64),
        ("first_failed_row", C.c_int64),
        ("kernel_ms", C.c_double),
        ("flops", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SurrogateStats(C.Structure):
    _fields_ = [
        ("net_evals", C.c_int64),
        ("exact_evals", C.c_int64),
        ("checks", C.c_int64),
        ("check_failures", C.c_int64),
        ("band_fallbacks", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Launch(C.Structure):
    _fields_ = [
        ("waves_per_group", C.c_int32),
        ("use_graph", C.c_int32),
        ("blocks", C.c_int32),
        ("sort_by_cost", C.c_int32),
        ("teams_per_block", C.c_int32),
    ]


class LaunchInfo(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32),
        ("waves_per_group", C.c_int32),
        ("teams", C.c_int32),
        ("blocks", C.c_int32),
        ("threads", C.c_int32),
    ]


# include/bqp.h BQP_KERNEL_*
KERNELS = {0: "none", 1: "fixed", 2: "team", 3: "pair"}

# every symbol include/bqp.h declares (checked by the ABI tests)
EXPORTS = [
    "bqp_default_config",
    "bqp_create",
    "bqp_destroy",
    "bqp_set_config",
    "bqp_set_launch",
    "bqp_last_error",
    "bqp_load_field",
    "bqp_gravity_accel_batch",
    "bqp_gravity_accel_batch_dev",
    "bqp_propagate_batch_final",
    "bqp_propagate_batch_final_dev",
    "bqp_propagate_batch_final_dev_noalloc",
    "bqp_propagate_batch_sampled",
    "bqp_get_tableau",
    "bqp_flops_per_rhs",
    "bqp_get_cost_model",
    "bqp_last_launch",
    "bqp_surrogate_param_count",
    "bqp_surrogate_init",
    "bqp_surrogate_get_params",
    "bqp_surrogate_set_params",
    "bqp_surrogate_dataset_dev",
    "bqp_surrogate_train_dev",
    "bqp_surrogate_loss_grad_dev",
    "bqp_surrogate_pack",
    "bqp_surrogate_infer_dev",
    "bqp_surrogate_set_check",
    "bqp_surrogate_last_stats",
]

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_lib = None


class BqpError(RuntimeError):
    """A bqp C-ABI call failed."""


def load(path: str | None = None):
    """Load libbqp.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise BqpError(
            f"{path} is missing: the HIP extension must be built (python -c 'import bqp_amd.build as b; b.build()')"
        )
    L = C.CDLL(path)
    L.bqp_default_config.argtypes = [C.POINTER(Config)]
    L.bqp_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.bqp_destroy.argtypes = [C.c_void_p]
    L.bqp_set_config.argtypes = [C.c_void_p, C.POINTER(Config)]
    L.bqp_set_launch.argtypes = [C.c_void_p, C.POINTER(Launch)]
    L.bqp_last_error.restype = C.c_char_p
    L.bqp_last_error.argtypes = [C.c_void_p]
    L.bqp_load_field.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_double]
    L.bqp_gravity_accel_batch.argtypes = [C.c_void_p, _dp, C.c_int64, _dp]
    L.bqp_gravity_accel_batch_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.bqp_propagate_batch_final.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int64, _dp, _i32p, _i32p,
                                            _i32p, C.POINTER(Stats)]
    L.bqp_propagate_batch_final_dev.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int64] + \
        [C.c_void_p] * 4 + [C.POINTER(Stats)]
    L.bqp_propagate_batch_final_dev_noalloc.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int64] + \
        [C.c_void_p] * 4 + [C.POINTER(Stats)]
    L.bqp_propagate_batch_sampled.argtypes = [C.c_void_p, _dp, C.c_double, _dp, C.c_int64, _dp, C.c_int64,
                                              _dp, _i32p, C.POINTER(Stats)]
    L.bqp_get_tableau.argtypes = [C.c_int32, C.POINTER(Tableau)]
    L.bqp_flops_per_rhs.restype = C.c_double
    L.bqp_flops_per_rhs.argtypes = [C.c_void_p]
    L.bqp_get_cost_model.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.bqp_last_launch.argtypes = [C.c_void_p, C.POINTER(LaunchInfo)]
    _fp = C.POINTER(C.c_float)
    vp = C.c_void_p
    L.bqp_surrogate_param_count.restype = C.c_int64
    L.bqp_surrogate_param_count.argtypes = []
    L.bqp_surrogate_init.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_uint64]
    L.bqp_surrogate_get_params.argtypes = [vp, _fp, _dp]
    L.bqp_surrogate_set_params.argtypes = [vp, _fp, _dp]
    L.bqp_surrogate_dataset_dev.argtypes = [vp, C.c_int64, C.c_uint64, vp, vp, vp]
    L.bqp_surrogate_train_dev.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_uint64,
                                          _fp]
    L.bqp_surrogate_loss_grad_dev.argtypes = [vp, vp, vp, C.c_int64, _dp, _fp]
    L.bqp_surrogate_pack.argtypes = [vp, C.c_int32, vp, C.c_int64]
    L.bqp_surrogate_infer_dev.argtypes = [vp, C.c_int32, vp, C.c_int64, vp, _dp]
    L.bqp_surrogate_set_check.argtypes = [vp, C.c_int32, C.c_double]
    L.bqp_surrogate_last_stats.argtypes = [vp, C.POINTER(SurrogateStats)]
    _lib = L
    return L
