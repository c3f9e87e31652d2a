"""ctypes mirror of ``bqp_tableau`` (bqp_amd/csrc/tableaus.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

MAX_STAGES = 13


class Tableau(C.Structure):
    _fields_ = [
        ("stages", C.c_int),
        ("order", C.c_int),
        ("order_err", C.c_int),
        ("fsal", C.c_int),
        ("has_e3", C.c_int),
        ("c", C.c_double * MAX_STAGES),
        ("a", (C.c_double * MAX_STAGES) * MAX_STAGES),
        ("b", C.c_double * MAX_STAGES),
        ("e", C.c_double * MAX_STAGES),
        ("e3", C.c_double * MAX_STAGES),
    ]


def to_dict(t: Tableau) -> dict:
    s = t.stages
    a = np.array([[t.a[i][j] for j in range(MAX_STAGES)] for i in range(MAX_STAGES)])
    return {
        "stages": s,
        "order": t.order,
        "order_err": t.order_err,
        "fsal": bool(t.fsal),
        "has_e3": bool(t.has_e3),
        "c": np.array(t.c[:s]),
        "a": a[:s, :s],
        "b": np.array(t.b[:s]),
        "e": np.array(t.e[:s]),
        "e3": np.array(t.e3[:s]),
    }
