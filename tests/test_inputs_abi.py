"""CPU tests of the host side: seeded inputs, element conversions, the field
loader, Python-boundary validation, and the C ABI's exported symbols (the
library is loaded, no compute call is made without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import bqp_amd
from bqp_amd import _native, inputs
from bqp_amd.elements import MU_EARTH, eci2kep, kep2eci

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "bqp_amd", "libbqp.so")


def test_kaula_field_prefix_and_determinism():
    C8, S8 = inputs.kaula_field(8)
    C70, S70 = inputs.kaula_field(70)
    k = inputs.tri_index(9, 0)
    assert np.array_equal(C8, C70[:k]) and np.array_equal(S8, S70[:k])
    C8b, _ = inputs.kaula_field(8)
    assert np.array_equal(C8, C8b)
    assert C8[0] == 1.0 and C8[1] == 0.0 and C8[2] == 0.0
    assert C8[inputs.tri_index(2, 0)] == inputs.C20
    # Kaula scale: std of degree-n coefficients ~ 1e-5 / n^2
    n = 60
    row = np.concatenate([C70[inputs.tri_index(n, 0):inputs.tri_index(n + 1, 0)],
                          S70[inputs.tri_index(n, 1):inputs.tri_index(n + 1, 0)]])
    assert 0.6 < np.std(row) / (1e-5 / n ** 2) < 1.4


def test_workload_shapes_and_ranges():
    w = inputs.workload("w1", n=2000)
    r = np.linalg.norm(w["y0"][:, :3], axis=1)
    assert w["y0"].shape == (2000, 6)
    assert np.all(r >= (6378137.0 + 400e3) * 0.98) and np.all(r <= (6378137.0 + 1200e3) * 1.02)
    assert np.all((w["bc"] >= 0.005) & (w["bc"] <= 0.02))
    w2 = inputs.workload("w2", n=5000)
    assert np.all(w2["tf"] >= 3600.0) and np.all(w2["tf"] <= 30 * 86400.0)
    frac = np.bincount(w2["orbit_class"], minlength=3) / 5000
    assert abs(frac[0] - 0.6) < 0.03 and abs(frac[1] - 0.25) < 0.03 and abs(frac[2] - 0.15) < 0.03
    w4 = inputs.workload("w4", n=4096)
    assert w4["generations"] * w4["population"] == 131072


def test_kep2eci_matches_oracle(orc):
    kep = inputs.leo_elements(50, seed=5)
    a = kep2eci(kep)
    b = orc.kep2eci(MU_EARTH, kep)
    assert np.max(np.abs(a - b) / np.linalg.norm(a, axis=1)[:, None]) < 1e-13


def test_eci2kep_round_trip_and_identities():
    kep = inputs.leo_elements(200, seed=9)
    s = kep2eci(kep)
    back = eci2kep(s, anomaly="mean")
    assert np.allclose(back[:, 0], kep[:, 0], rtol=1e-12)
    assert np.allclose(back[:, 1], kep[:, 1], atol=1e-10)
    assert np.allclose(back[:, 2], kep[:, 2], atol=1e-10)
    r = np.linalg.norm(s[:, :3], axis=1)
    v2 = np.sum(s[:, 3:] ** 2, axis=1)
    assert np.allclose(v2, MU_EARTH * (2 / r - 1 / kep[:, 0]), rtol=1e-12)  # vis-viva
    h = np.linalg.norm(np.cross(s[:, :3], s[:, 3:]), axis=1)
    assert np.allclose(h, np.sqrt(MU_EARTH * kep[:, 0] * (1 - kep[:, 1] ** 2)), rtol=1e-12)


def test_reference_shaped_kep_to_eci():
    s = bqp_amd.kep_to_eci([7000.0, 0.001, 51.6, 10.0, 20.0, 30.0])
    k = bqp_amd.eci_to_kep(s)
    assert np.allclose(k, [7000.0, 0.001, 51.6, 10.0, 20.0, 30.0], rtol=1e-9, atol=1e-7)
    with pytest.raises(ValueError):
        bqp_amd.kep_to_eci([7000.0, 1.5, 0, 0, 0, 0])
    with pytest.raises(ValueError):
        bqp_amd.kep_to_eci([1, 2, 3])


def test_goce_field_loader():
    f = bqp_amd.goce(15)
    assert f.nmax == 15 and f.cbar[0] == 1.0
    assert f.cbar[inputs.tri_index(2, 0)] == pytest.approx(-4.84165301385e-4, rel=1e-12)
    assert f.mu == 3.986004415e14 and f.radius == 6378136.46
    t = f.truncated(5)
    assert t.nmax == 5 and t.cbar.size == inputs.tri_index(6, 0)


def test_icgem_gfc_parser(tmp_path):
    p = tmp_path / "f.gfc"
    p.write_text("product_type gravity_field\nearth_gravity_constant 3.986004415E+14\nradius 6378136.3\n"
                 "max_degree 3\nnorm fully_normalized\nend_of_head ====\n"
                 "gfc 2 0 -4.84D-04 0.0 1e-12 1e-12\ngfc 2 2 2.4e-06 -1.4e-06 0 0\ngfc 3 1 2.0e-06 2.5e-07 0 0\n")
    f = bqp_amd.load_icgem(str(p))
    assert f.nmax == 3 and f.radius == 6378136.3
    assert f.cbar[inputs.tri_index(2, 0)] == -4.84e-4 and f.sbar[inputs.tri_index(2, 2)] == -1.4e-6


def test_reference_defaults_name_out_of_scope_models():
    """The reference-shaped calls keep the reference's keyword defaults
    (dust_transfer/src/lib.rs:238-241, 364-366: srp, sun, moon on, JB2008,
    Vern7).  Those models are outside this propagator, so relying on any one
    of them raises ValueError before any device work; the in-scope set
    passes the model checks."""
    import inspect

    for fn in (bqp_amd.propagate, bqp_amd.propagate_final, bqp_amd.propagate_batch_final):
        kw = {k: v.default for k, v in inspect.signature(fn).parameters.items()
              if v.kind == inspect.Parameter.KEYWORD_ONLY}
        assert kw["srp"] is True and kw["sun"] is True and kw["moon"] is True
        assert kw["atm_model"] == 7 and kw["method"] == "vern7" and kw["gravity_order"] == 5
        assert kw["drag"] is True and kw["tol"] == 1e-8 and kw["dt_max_s"] == 300.0
    s = bqp_amd.kep_to_eci([7000.0, 0.001, 51.6, 10.0, 20.0, 30.0])
    ok = bqp_amd.IN_SCOPE
    with pytest.raises(ValueError, match="srp, sun and moon"):
        bqp_amd.propagate_final(s, 2451545.0, 600.0)
    with pytest.raises(ValueError, match="srp, sun and moon"):
        bqp_amd.propagate(s, 2451545.0, 600.0)
    with pytest.raises(ValueError, match="srp, sun and moon"):
        bqp_amd.propagate_batch_final(s[None, :], [600.0])
    for key in ("sun", "moon"):
        with pytest.raises(ValueError, match="srp, sun and moon"):
            bqp_amd.propagate_final(s, 2451545.0, 600.0, **{**ok, key: True})
    with pytest.raises(ValueError, match="atm_model 7"):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, **{k: v for k, v in ok.items() if k != "atm_model"})
    with pytest.raises(ValueError, match="vern7"):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, **{k: v for k, v in ok.items() if k != "method"})
    # drag off: the atmosphere model is not consulted (as in the reference)
    field, kw, _ = bqp_amd._model(5, False, False, False, False, 7, 0.01, 2.2, 1.3, 1e-8, 300.0, "dop853",
                                  2451545.0, ground=True)
    assert kw["drag"] == "none" and field.nmax == 5


def test_python_boundary_validation_before_device():
    s = bqp_amd.kep_to_eci([7000.0, 0.001, 51.6, 10.0, 20.0, 30.0])
    ok = bqp_amd.IN_SCOPE
    with pytest.raises(ValueError):
        bqp_amd.propagate_final(s, 2451545.0, -1.0, **ok)
    with pytest.raises(ValueError):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, **{**ok, "srp": True})
    with pytest.raises(ValueError):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, gravity_order=16, **ok)
    with pytest.raises(ValueError):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, **{**ok, "method": "vern7"})
    with pytest.raises(ValueError):
        bqp_amd.propagate_final(s, 2451545.0, 600.0, **{**ok, "atm_model": 7})
    with pytest.raises(ValueError):
        bqp_amd.propagate_final([1, 2, np.nan, 4, 5, 6], 2451545.0, 600.0, **ok)
    with pytest.raises(ValueError):
        bqp_amd.propagate(s, 2451545.0, 600.0, times_s=[100.0, 50.0], **ok)
    # not a bound ellipse (above escape speed; radial, zero angular momentum)
    fast = s.copy()
    fast[3:] *= 1.5
    radial = np.array([7000.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    for bad in (fast, radial):
        with pytest.raises(ValueError, match="bound, non-degenerate"):
            bqp_amd.propagate_final(bad, 2451545.0, 600.0)
        with pytest.raises(ValueError, match="bound, non-degenerate"):
            bqp_amd.propagate(bad, 2451545.0, 600.0)


def _header_functions():
    with open(os.path.join(ROOT, "include", "bqp.h")) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(bqp_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_exactly_the_bound_symbols():
    assert _header_functions() == sorted(_native.EXPORTS)


def test_cost_model_cells_match_the_header():
    with open(os.path.join(ROOT, "include", "bqp.h")) as fh:
        txt = fh.read()
    defines = dict(re.findall(r"#define (BQP_COST_[A-Z_]+) (\d+)", txt))
    assert int(defines["BQP_COST_ALT_BINS"]) == _native.COST_ALT_BINS
    assert int(defines["BQP_COST_INC_BINS"]) == _native.COST_INC_BINS


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbqp.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for name in _header_functions():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbqp.so not built")
def test_host_only_entry_points_without_gpu():
    L = _native.load()
    cfg = _native.Config()
    assert L.bqp_default_config(ctypes.byref(cfg)) == 0
    assert cfg.method == 2 and cfg.rtol == 1e-12
    assert L.bqp_default_config(None) == _native.E_INVALID
    # null-context calls fail cleanly with E_INVALID, no device touched
    assert L.bqp_destroy(None) == 0
    assert L.bqp_load_field(None, 2, 2, None, None, 1.0, 1.0) == _native.E_INVALID
    assert L.bqp_propagate_batch_final(None, None, None, None, None, 0, None, None, None, None, None) == \
        _native.E_INVALID
    assert L.bqp_last_error(None) == b"null context"
    assert L.bqp_get_cost_model(None, None, None, None) == _native.E_INVALID
    assert L.bqp_last_launch(None, None) == _native.E_INVALID
    info = _native.LaunchInfo()
    assert L.bqp_last_launch(None, ctypes.byref(info)) == _native.E_INVALID


def test_sampled_call_returns_only_the_samples_a_failed_arc_reached():
    grid = [0.0, 600.0, 1200.0, 3600.0]
    st = np.arange(24, dtype=float).reshape(4, 6)
    t, y = bqp_amd._reached_samples(grid, st, 0)
    assert list(t) == grid and np.array_equal(y, st)
    failed = st.copy()
    failed[2:] = np.nan  # the kernel NaN-fills from the first unreached sample
    t, y = bqp_amd._reached_samples(grid, failed, 1)
    assert list(t) == [0.0, 600.0] and np.array_equal(y, st[:2])
    failed[:] = np.nan
    t, y = bqp_amd._reached_samples(grid, failed, 5)
    assert len(t) == 0 and y.shape == (0, 6)


def test_batch_min_radius_precheck():
    """Rows are refused up front when the radius or the osculating perigee is
    below R + 100 km (the reference's batch, session.rs:1983-1995)."""
    r_min = 6378.137 + 100.0
    good = bqp_amd.kep_to_eci([6878.0, 0.001, 51.6, 0.0, 0.0, 0.0])
    low_perigee = bqp_amd.kep_to_eci([6600.0, 0.03, 51.6, 0.0, 0.0, 180.0])  # at apogee 6798 km, perigee 6402 km
    low_now = bqp_amd.kep_to_eci([6450.0, 0.0, 51.6, 0.0, 0.0, 0.0])
    ok = bqp_amd._clears_min_radius(np.array([good, low_perigee, low_now]), r_min)
    assert list(ok) == [True, False, False]


class _StubBuf:
    """A device-buffer stand-in: shape, dtype and a data_ptr() that must
    never be reached (the contract check runs before any native call)."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)

    def data_ptr(self):
        raise AssertionError("native call reached with a buffer that breaks the contract")


def test_device_entry_point_rejects_bad_buffers_before_native_call():
    """propagate_batch_final_dev takes raw device pointers; the kernel reads
    y0/t0/tf/bc and writes yf/status/steps/evals as (n, 6) float64 and (n,)
    float64 / int32 with no bounds of its own, so a wrong shape or dtype must
    raise ValueError on the host.  The context is never touched: the check
    runs before the ctypes call."""
    from bqp_amd.propagator import Propagator

    p = Propagator.__new__(Propagator)  # no context: any native call would fail
    n = 8
    good = dict(y0=_StubBuf((n, 6), np.float64), t0=_StubBuf((n,), np.float64), tf=_StubBuf((n,), np.float64),
                bc=_StubBuf((n,), np.float64), yf=_StubBuf((n, 6), np.float64), status=_StubBuf((n,), np.int32),
                steps=_StubBuf((n,), np.int32), evals=_StubBuf((n,), np.int32))
    bad_cases = [
        ("y0", _StubBuf((n * 6,), np.float64)),       # flat states
        ("y0", _StubBuf((n, 6), np.float32)),
        ("t0", _StubBuf((1,), np.float64)),           # a scalar time as a 1-element buffer
        ("tf", _StubBuf((n,), np.int64)),
        ("tf", _StubBuf((n + 1,), np.float64)),
        ("bc", _StubBuf((n - 1,), np.float64)),
        ("yf", _StubBuf((n, 3), np.float64)),
        ("status", _StubBuf((n,), np.int64)),
        ("steps", _StubBuf((n, 1), np.int32)),
        ("evals", _StubBuf((n,), np.float32)),
        ("t0", np.zeros(n)),                          # a host array, not a device buffer
    ]
    for name, buf in bad_cases:
        kw = {**good, name: buf}
        with pytest.raises(ValueError, match=name):
            p.propagate_batch_final_dev(**kw)
    with pytest.raises(ValueError, match="t0"):
        p.propagate_batch_final_dev(good["y0"], None, good["tf"])


def test_epoch_offsets_reproduce_each_rows_frame_angle():
    """Row i of a batch starts at t = offset[i] of a frame turning as
    theta0 + omega t; there its angle equals gmst(epoch_jd[i])."""
    ep = np.array([2451545.0, 2451545.25, 2451547.7, 2451544.9, 2460000.5])
    theta0, off = bqp_amd.epoch_offsets(ep)
    assert theta0 == bqp_amd.gmst(ep[0]) and off[0] == 0.0
    assert np.all(off >= 0.0) and np.all(off < 2 * np.pi / bqp_amd.OMEGA_EARTH)
    for e, o in zip(ep, off):
        d = (theta0 + bqp_amd.OMEGA_EARTH * o - bqp_amd.gmst(e) + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) < 1e-12


def test_per_row_batch_arguments_are_validated():
    s = np.array([bqp_amd.kep_to_eci([7000.0, 0.001, 51.6, 10.0, 20.0, 30.0])] * 3)
    ok = bqp_amd.IN_SCOPE
    with pytest.raises(ValueError, match="one value per row"):
        bqp_amd.propagate_batch_final(s, 600.0, epoch_jd=[2451545.0, 2451546.0], **ok)
    with pytest.raises(ValueError, match="every row"):
        bqp_amd.propagate_batch_final(s, 600.0, am_ratio=[0.01, -0.01, 0.01], **ok)
    with pytest.raises(ValueError, match="every row"):
        bqp_amd.propagate_batch_final(s, 600.0, cd=[2.2, np.nan, 2.2], **ok)
    with pytest.raises(ValueError, match="finite"):
        bqp_amd.propagate_batch_final(s, 600.0, epoch_jd=[2451545.0, np.inf, 2451545.0], **ok)
