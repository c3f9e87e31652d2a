"""Batch wire format (NPZ and Arrow IPC) round trips, NaN rows, errors."""
import numpy as np
import pytest

from bqp_amd import batchio, inputs


@pytest.mark.parametrize("ext", [".npz", ".arrow"])
def test_inputs_round_trip(tmp_path, ext):
    w = inputs.workload("w2", n=300)
    p = str(tmp_path / ("batch" + ext))
    batchio.save_inputs(p, w["y0"], 0.0, w["tf"], w["bc"] if "bc" in w else None)
    y0, t0, tf, bc = batchio.load_inputs(p)
    assert np.array_equal(y0, w["y0"]) and np.array_equal(tf, w["tf"]) and np.all(t0 == 0.0)
    if "bc" in w:
        assert np.array_equal(bc, w["bc"])


@pytest.mark.parametrize("ext", [".npz", ".arrow"])
def test_results_round_trip_keeps_nan_rows(tmp_path, ext):
    rng = np.random.default_rng(0)
    yf = rng.normal(size=(50, 6)) * 7e6
    status = np.zeros(50, np.int32)
    yf[[3, 17]] = np.nan
    status[[3, 17]] = [5, 6]
    steps = rng.integers(0, 1000, 50).astype(np.int32)
    p = str(tmp_path / ("res" + ext))
    batchio.save_results(p, yf, status, steps=steps)
    a, s, st, ev = batchio.load_results(p)
    assert np.array_equal(a, yf, equal_nan=True) and np.array_equal(s, status) and np.array_equal(st, steps)
    assert ev is None


def test_arrow_columns_readable_without_package(tmp_path):
    import pyarrow as pa

    p = str(tmp_path / "res.arrow")
    batchio.save_results(p, np.ones((4, 6)), np.zeros(4, np.int32))
    with pa.memory_map(p, "r") as src:
        t = pa.ipc.open_file(src).read_all()
    assert t.column_names == ["yf_x", "yf_y", "yf_z", "yf_vx", "yf_vy", "yf_vz", "status"]
    assert t.num_rows == 4


def test_ragged_columns_rejected(tmp_path):
    with pytest.raises(ValueError):
        batchio.save_results(str(tmp_path / "bad.npz"), np.ones((4, 6)), np.zeros(3, np.int32))
