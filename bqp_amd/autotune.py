"""Launch-shape autotuner for the persistent propagation kernel.

Tunables (``bqp_launch``): G, the number of waves that share one group of 64
trajectories (1, 2, 4, 8, 16; the workgroup is 64·G threads and holds 64
trajectories), the persistent grid size, the longest-first hand-out order,
and -- for FP64, split and Encke -- the pair kernel (two groups per 8-wave
workgroup, ``waves_per_group=8, teams_per_block=2``) and, for fields of
degree >= 30, the team kernel with 1, 2 or 4 groups per 16-wave workgroup
(``teams_per_block``), whose waves move to the groups still running.  The tuner runs a short pilot of the real batch — every row's span
scaled by ``pilot_fraction`` so the spread of row lengths (and hence the
queue behaviour) is kept — for each candidate, keeps the fastest by kernel
time, caches it per workload signature, and sets it on the propagator.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .device import DeviceArray

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "bqp_amd", "launch.json")


def signature(prop, n: int, spans: np.ndarray) -> str:
    c = prop.config
    spread = float(np.log10(max(spans.max(), 1.0) / max(spans.min(), 1.0))) if spans.size else 0.0
    return (f"m{c.method}-p{c.precision}-d{c.drag}-N{prop.field.nmax}-n{int(math.log2(max(n, 1)))}"
            f"-s{int(round(spread))}")


def _load_cache(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def autotune(prop, y0, t0, tf, bc=None, *, pilot_fraction: float = 0.01, min_pilot_s: float = 60.0,
             candidates=None, cache_path: str | None = CACHE, use_cache: bool = True) -> dict:
    """Pick and set the fastest launch shape for this batch; returns a report."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
    n = y0.shape[0]
    t0 = np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,))
    tf = np.broadcast_to(np.asarray(tf, dtype=np.float64), (n,))
    spans = tf - t0
    key = signature(prop, n, spans)
    if use_cache and cache_path:
        hit = _load_cache(cache_path).get(key)
        if hit:
            prop.set_launch(**hit["launch"])
            return {"signature": key, "cached": True, **hit}
    if candidates is None:
        candidates = [{"waves_per_group": g, "sort_by_cost": s} for g in (1, 2, 4, 8, 16) for s in (1, 0)]
        if prop.config.precision in (0, 1):
            candidates.append({"waves_per_group": 8, "teams_per_block": 2, "sort_by_cost": 1})
        if prop.field.nmax >= 30 and prop.config.precision == 0:
            candidates += [{"waves_per_group": 0, "teams_per_block": t, "sort_by_cost": 1} for t in (1, 2, 4)]
    pilot_tf = t0 + np.maximum(spans * pilot_fraction, np.minimum(spans, min_pilot_s))
    d_y0 = DeviceArray.from_numpy(y0)
    d_t0 = DeviceArray.from_numpy(np.ascontiguousarray(t0))
    d_tf = DeviceArray.from_numpy(np.ascontiguousarray(pilot_tf))
    d_bc = DeviceArray.from_numpy(np.broadcast_to(np.asarray(bc, dtype=np.float64), (n,))) if bc is not None else None
    d_yf = DeviceArray((n, 6))
    d_st = DeviceArray((n,), np.int32)
    trials = []
    saved = getattr(prop, "launch", None)
    for cand in candidates:
        prop.set_launch(**cand)
        _, _, st = prop.propagate_batch_final_dev(d_y0, d_t0, d_tf, d_bc, d_yf, d_st)
        trials.append({"launch": dict(cand), "kernel_ms": st["kernel_ms"]})
    best = min(trials, key=lambda r: r["kernel_ms"])
    prop.set_launch(**best["launch"])
    report = {"signature": key, "cached": False, "launch": best["launch"], "pilot_kernel_ms": best["kernel_ms"],
              "trials": trials, "previous": saved}
    if cache_path:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        data = _load_cache(cache_path)
        data[key] = {"launch": best["launch"], "pilot_kernel_ms": best["kernel_ms"]}
        with open(cache_path, "w") as fh:
            json.dump(data, fh, indent=1)
    return report
