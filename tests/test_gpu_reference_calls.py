"""The reference-shaped Python entry points (``propagate_final``,
``propagate``, ``propagate_batch_final``; km, km/s, s, Julian dates) on the
GPU, against the oracle configured as those calls configure the device: the
GOCE field truncated to ``gravity_order``, the reference's single-H
exponential drag (``atm_model=1``) with Cd*A/m = cd * am_ratio, the body frame
at GMST(epoch), Encke deviations rebased every 5,400 s, DOP853 with absolute
error control at ``tol`` km and a 300 s step cap, ground impact checked.
Bar: 1 cm."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import IN_SCOPE, OMEGA_EARTH, R_EARTH
from gpu_util import announce

pytestmark = pytest.mark.gpu

EPOCH = 2459000.5


def _oracle(orc, s_km, span, gravity_order=5, tol=1e-8, am_ratio=0.01, cd=2.2, ground=True):
    field = bqp_amd.goce(gravity_order)
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    m = orc.make_model(f, omega=OMEGA_EARTH, theta0=bqp_amd.gmst(EPOCH), drag="compat", r_atm=R_EARTH,
                       r_min=(R_EARTH + 100e3) if ground else 0.0, formulation="encke",
                       ground_flattening=bqp_amd.WGS84_FLATTENING if ground else 0.0)
    i = orc.make_integ(method="dop853", control="absolute", eps=tol * 1e3, h_max=300.0)
    y0 = np.asarray(s_km, dtype=np.float64).reshape(-1, 6) * 1e3
    yo, so, _, _ = orc.propagate_batch(m, i, y0, 0.0, span, np.full(y0.shape[0], am_ratio * cd))
    return yo * 1e-3, so


def test_propagate_final_matches_oracle(orc):
    announce("propagate_final (km, GOCE d/o 5, exponential drag, Encke, DOP853) vs oracle")
    s = bqp_amd.kep_to_eci([6878.0, 0.001, 51.6, 10.0, 20.0, 30.0])
    yk = bqp_amd.propagate_final(s, EPOCH, 6 * 3600.0, **IN_SCOPE)
    yo, so = _oracle(orc, s, 6 * 3600.0)
    assert so[0] == 0
    assert np.linalg.norm(yk[:3] - yo[0, :3]) * 1e3 < 1e-2


def test_propagate_samples_match_final_and_oracle(orc):
    announce("propagate (sampled dict) vs propagate_final and the oracle")
    s = bqp_amd.kep_to_eci([7200.0, 0.01, 98.0, 40.0, 60.0, 0.0])
    tof = 4 * 3600.0
    out = bqp_amd.propagate(s, EPOCH, tof, times_s=[600.0, 5400.0, 9000.0], **IN_SCOPE)
    assert set(out) == {"times_s", "states_eci", "completed", "terminal_event", "max_steps_exceeded", "steps",
                        "rhs_evals", "wall_us"}
    assert out["completed"] and not out["max_steps_exceeded"] and out["terminal_event"] == ""
    assert list(out["times_s"]) == [0.0, 600.0, 5400.0, 9000.0, tof]
    assert np.array_equal(out["states_eci"][0], s)
    assert out["steps"] > 0 and out["rhs_evals"] >= out["steps"]
    yo, so = _oracle(orc, s, tof, ground=False)
    assert so[0] == 0
    assert np.linalg.norm(out["states_eci"][-1, :3] - yo[0, :3]) * 1e3 < 1e-2
    yf = bqp_amd.propagate_final(s, EPOCH, tof, **IN_SCOPE)
    assert np.linalg.norm(out["states_eci"][-1, :3] - yf[:3]) * 1e3 < 1e-2


def test_propagate_batch_final_matches_oracle_and_flags_impact(orc):
    announce("propagate_batch_final (km) vs oracle; a decaying row is flagged, not fatal")
    kep = np.array([[6778.0, 0.001, 51.6, 0.0, 0.0, 0.0], [7000.0, 0.002, 28.5, 100.0, 30.0, 45.0],
                    [26560.0, 0.01, 55.0, 200.0, 10.0, 300.0], [6500.0, 0.02, 10.0, 0.0, 0.0, 180.0]])
    s = np.array([bqp_amd.kep_to_eci(k) for k in kep])
    tof = np.array([3 * 3600.0, 5 * 3600.0, 12 * 3600.0, 3 * 3600.0])
    yk, st = bqp_amd.propagate_batch_final(s, tof, epoch_jd=EPOCH, **IN_SCOPE)
    for k in range(3):
        yo_k, so_k = _oracle(orc, s[k:k + 1], tof[k])
        assert st[k] == 0 and so_k[0] == 0
        assert np.linalg.norm(yk[k, :3] - yo_k[0, :3]) * 1e3 < 1e-2
    # perigee 6500 km x (1 - 0.02) is below R + 100 km: refused up front,
    # NaN-filled; the single-arc call reaches the ground event in flight
    assert st[3] == 6 and np.all(np.isnan(yk[3]))
    with pytest.raises(ValueError, match="ground"):
        bqp_amd.propagate_final(s[3], EPOCH, tof[3], **IN_SCOPE)


def test_ground_event_over_the_ellipsoid_matches_oracle(orc):
    announce("ground event 100 km above WGS84: polar state at 110 km clears it, the sphere would not")
    R, fl = R_EARTH, bqp_amd.WGS84_FLATTENING
    field = bqp_amd.kaula(4)
    rp = R * (1.0 - fl) + 110e3
    polar = np.array([[0.0, 0.0, rp, np.sqrt(field.mu / rp), 0.0, 0.0]])
    for flat, want in ((fl, 0), (0.0, 6)):
        with bqp_amd.Propagator(field, r_min=R + 100e3, ground_flattening=flat) as p:
            _, st, _ = p.propagate_batch_final(polar, 0.0, 60.0)
        f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
        m = orc.make_model(f, r_atm=R, r_min=R + 100e3, ground_flattening=flat)
        i = orc.make_integ(method="rk87", rtol=1e-12, canonical=(field.radius, field.mu))
        _, so, _, _ = orc.propagate_batch(m, i, polar, 0.0, 60.0)
        assert int(st[0]) == want and int(so[0]) == want



def test_batch_rows_with_own_epochs_and_ballistics_equal_separate_calls():
    """The reference's ragged batch carries an epoch and ballistic
    coefficients per row (VariableFinalBatchRequest, LY/src/session.rs:
    2028-2035).  Rows with different epochs (spread over three days) and
    different A/m and Cd in one call agree with one call per row at that
    row's epoch and ballistics; the per-row clock offsets only move each
    row's time origin."""
    announce("propagate_batch_final: per-row epochs and ballistics vs one call per row")
    kep = np.array([[6778.0, 0.001, 51.6, 0.0, 0.0, 0.0], [7000.0, 0.002, 28.5, 100.0, 30.0, 45.0],
                    [6900.0, 0.01, 97.8, 200.0, 10.0, 300.0], [7100.0, 0.005, 63.4, 10.0, 80.0, 120.0]])
    s = np.array([bqp_amd.kep_to_eci(k) for k in kep])
    tof = np.array([3 * 3600.0, 5 * 3600.0, 4 * 3600.0, 2 * 3600.0])
    epochs = np.array([2460000.5, 2460001.37, 2459998.81, 2460002.05])
    am = np.array([0.01, 0.02, 0.005, 0.015])
    cd = np.array([2.2, 2.0, 2.4, 2.3])
    yk, st = bqp_amd.propagate_batch_final(s, tof, epoch_jd=epochs, am_ratio=am, cd=cd, **IN_SCOPE)
    assert np.all(st == 0)
    for i in range(4):
        yi, si = bqp_amd.propagate_batch_final(s[i:i + 1], tof[i:i + 1], epoch_jd=epochs[i], am_ratio=am[i],
                                               cd=cd[i], **IN_SCOPE)
        assert si[0] == 0
        assert np.linalg.norm(yk[i, :3] - yi[0, :3]) * 1e3 < 1e-3
    # the epoch matters: the same rows at one common epoch end elsewhere
    y0, _ = bqp_amd.propagate_batch_final(s, tof, epoch_jd=epochs[0], am_ratio=am, cd=cd, **IN_SCOPE)
    assert np.all(np.linalg.norm(yk[1:, :3] - y0[1:, :3], axis=1) * 1e3 > 1e-3)
