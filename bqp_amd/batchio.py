"""Wire format of batch inputs and results: NPZ and Arrow IPC.

The reference hands batches across its Python boundary as numpy arrays
(DT/src/lib.rs:200-438; the batch request of LY/src/batch.rs:120-133 holds
`[n x 6]` states, per-row times and per-row ballistics).  Here the same
columns travel as a file:

  inputs : y0 (n, 6) f64 [m, m/s, inertial], t0 (n,) f64 [s], tf (n,) f64 [s],
           bc (n,) f64 [Cd*A/m, m^2/kg; optional]
  results: yf (n, 6) f64, status (n,) i32 (row codes 0..6), steps (n,) i32,
           evals (n,) i32 (optional)

NPZ keeps the arrays as they are (``numpy.load`` with ``allow_pickle=False``
reads it back).  Arrow IPC stores one row per trajectory with columns
``x y z vx vy vz`` (or ``yf_x ...``) plus the scalar columns, so other tools
(pandas, DuckDB, Spark) can read the batch without this package.
"""
from __future__ import annotations

import numpy as np

STATE_COLS = ("x", "y", "z", "vx", "vy", "vz")


def _check_rows(arrays: dict) -> int:
    n = None
    for k, v in arrays.items():
        if v is None:
            continue
        m = v.shape[0]
        if n is None:
            n = m
        elif m != n:
            raise ValueError(f"column {k!r} has {m} rows, expected {n}")
    return 0 if n is None else n


def _inputs(y0, t0, tf, bc):
    y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 6)
    n = y0.shape[0]
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(tf, dtype=np.float64), (n,)))
    bc = None if bc is None else np.ascontiguousarray(np.broadcast_to(np.asarray(bc, dtype=np.float64), (n,)))
    return y0, t0, tf, bc


def save_inputs(path: str, y0, t0, tf, bc=None):
    """Write a batch request (.npz or .arrow / .feather by extension)."""
    y0, t0, tf, bc = _inputs(y0, t0, tf, bc)
    cols = {"t0": t0, "tf": tf}
    if bc is not None:
        cols["bc"] = bc
    _write(path, "", y0, cols)


def load_inputs(path: str):
    """Read a batch request: (y0, t0, tf, bc or None)."""
    y0, cols = _read(path, "")
    for k in ("t0", "tf"):
        if k not in cols:
            raise ValueError(f"{path}: missing column {k!r}")
    return y0, cols["t0"], cols["tf"], cols.get("bc")


def save_results(path: str, yf, status, steps=None, evals=None):
    """Write batch results (failed rows keep their NaN states and status)."""
    yf = np.ascontiguousarray(yf, dtype=np.float64).reshape(-1, 6)
    cols = {"status": np.ascontiguousarray(status, dtype=np.int32)}
    if steps is not None:
        cols["steps"] = np.ascontiguousarray(steps, dtype=np.int32)
    if evals is not None:
        cols["evals"] = np.ascontiguousarray(evals, dtype=np.int32)
    _write(path, "yf_", yf, cols)


def load_results(path: str):
    """Read batch results: (yf, status, steps or None, evals or None)."""
    yf, cols = _read(path, "yf_")
    if "status" not in cols:
        raise ValueError(f"{path}: missing column 'status'")
    return yf, cols["status"], cols.get("steps"), cols.get("evals")


def _is_arrow(path: str) -> bool:
    return path.endswith((".arrow", ".feather", ".ipc"))


def _write(path: str, prefix: str, states: np.ndarray, cols: dict):
    _check_rows({"states": states, **cols})
    if _is_arrow(path):
        import pyarrow as pa

        names = [prefix + c for c in STATE_COLS]
        arrays = [pa.array(states[:, i]) for i in range(6)] + [pa.array(v) for v in cols.values()]
        table = pa.table(arrays, names=names + list(cols))
        with pa.OSFile(path, "wb") as sink, pa.ipc.new_file(sink, table.schema) as w:
            w.write_table(table)
    else:
        np.savez(path, **{prefix + "states" if prefix else "y0": states}, **cols)


def _read(path: str, prefix: str):
    if _is_arrow(path):
        import pyarrow as pa

        with pa.memory_map(path, "r") as src:
            table = pa.ipc.open_file(src).read_all()
        names = [prefix + c for c in STATE_COLS]
        missing = [c for c in names if c not in table.column_names]
        if missing:
            raise ValueError(f"{path}: missing state columns {missing}")
        states = np.stack([table.column(c).to_numpy() for c in names], axis=1).astype(np.float64)
        cols = {c: table.column(c).to_numpy() for c in table.column_names if c not in names}
    else:
        with np.load(path, allow_pickle=False) as z:
            key = prefix + "states" if prefix else "y0"
            if key not in z:
                raise ValueError(f"{path}: missing array {key!r}")
            states = z[key].astype(np.float64).reshape(-1, 6)
            cols = {k: z[k] for k in z.files if k != key}
    _check_rows({"states": states, **cols})
    return states, cols
