"""Pins the oracle's integrator (orc_propagate) with closed forms and invariants.

* Two-body arcs against the Kepler closed form, for every method.
* Observed convergence order with fixed steps (h_min = h_max = h0, loose
  tolerance so every step is accepted) on the two-body problem.
* Jacobi integral E - w.(r x v) of a uniformly rotating conservative field.
* Tolerance-refinement convergence and cross-method agreement.
* Controller behaviours restated from LY/src/odesolve/basic_tests.rs:394-500:
  reject/h_min/max-steps statuses, NaN-filled failed rows.
* Drag dissipates energy.
"""
import numpy as np
import pytest

from bqp_amd import inputs
from bqp_amd.elements import MU_EARTH, R_EARTH, kep2eci, solve_kepler

DAY = 86400.0


def point_mass_field(orc, N=2):
    C = np.zeros((N + 1) * (N + 2) // 2)
    C[0] = 1.0
    return orc.Field(C, np.zeros_like(C), N, MU_EARTH, R_EARTH)


def kepler_state(kep, t):
    k = kep.copy()
    n = np.sqrt(MU_EARTH / k[0] ** 3)
    k[5] = k[5] + n * t
    return kep2eci(k)


@pytest.mark.parametrize("method,rtol", [("dopri5", 1e-11), ("dop853", 1e-12), ("rk87", 1e-12)])
def test_two_body_closed_form(orc, method, rtol):
    f = point_mass_field(orc)
    model = orc.make_model(f, omega=0.0)
    integ = orc.make_integ(method=method, rtol=rtol)
    kep = inputs.leo_elements(8, seed=11)
    y0 = kep2eci(kep)
    T = 0.5 * DAY
    yf, st, steps, evals = orc.propagate_batch(model, integ, y0, 0.0, T)
    assert np.all(st == 0)
    ref = np.array([kepler_state(k, T) for k in kep])
    err = np.linalg.norm(yf[:, :3] - ref[:, :3], axis=1)
    assert err.max() < (0.1 if method == "dopri5" else 1e-3), err.max()


@pytest.mark.parametrize("method,order", [("dopri5", 5), ("dop853", 8), ("rk87", 8)])
def test_fixed_step_convergence_order(orc, method, order):
    f = point_mass_field(orc)
    model = orc.make_model(f, omega=0.0)
    kep = np.array([R_EARTH + 700e3, 0.05, 0.9, 0.3, 1.1, 0.2])
    y0 = kep2eci(kep)
    T = 2400.0
    ref = kepler_state(kep, T)
    errs = []
    hs = [120.0, 60.0] if order == 5 else [240.0, 120.0]
    for h in hs:
        integ = orc.make_integ(method=method, control="absolute", eps=1e30, h0=h, h_min=h, h_max=h)
        yf, st = orc.propagate(model, integ, y0, 0.0, T)
        assert st.status == 0
        errs.append(np.linalg.norm(yf[:3] - ref[:3]))
    p = np.log2(errs[0] / errs[1])
    assert order - 0.6 < p < order + 1.5, (errs, p)


def test_jacobi_integral_rotating_field(orc):
    """Without drag the Jacobi integral E - w (x vy - y vx) is conserved in a
    field rotating uniformly about z."""
    field = inputs.kaula_field(20)
    f = orc.Field(field[0], field[1], 20, MU_EARTH, R_EARTH)
    model = orc.make_model(f, omega=7.292115e-5)
    integ = orc.make_integ(method="rk87", rtol=1e-13)
    y0 = inputs.workload("w0", n=4)["y0"]
    yf, st, _, _ = orc.propagate_batch(model, integ, y0, 0.0, 0.25 * DAY)
    assert np.all(st == 0)
    from potential_ref import potential_rotating

    th0, th1 = 0.0, 7.292115e-5 * 0.25 * DAY
    J0 = [0.5 * y @ y - potential_rotating(field, 20, y_[:3], th0) - 7.292115e-5 * (y_[0] * y_[4] - y_[1] * y_[3])
          for y, y_ in zip(y0[:, 3:], y0)]
    J1 = [0.5 * y @ y - potential_rotating(field, 20, y_[:3], th1) - 7.292115e-5 * (y_[0] * y_[4] - y_[1] * y_[3])
          for y, y_ in zip(yf[:, 3:], yf)]
    rel = np.max(np.abs(np.array(J1) - np.array(J0)) / np.abs(np.array(J0)))
    assert rel < 1e-11, rel


def test_tolerance_refinement_converges(orc):
    field = inputs.kaula_field(12)
    f = orc.Field(field[0], field[1], 12, MU_EARTH, R_EARTH)
    model = orc.make_model(f)
    y0 = inputs.workload("w0", n=4)["y0"]
    finals = []
    for rtol in (1e-10, 1e-11, 1e-12, 1e-13):
        integ = orc.make_integ(method="rk87", rtol=rtol)
        yf, st, _, _ = orc.propagate_batch(model, integ, y0, 0.0, 0.5 * DAY)
        assert np.all(st == 0)
        finals.append(yf)
    d = [np.max(np.linalg.norm(finals[i][:, :3] - finals[-1][:, :3], axis=1)) for i in range(3)]
    assert d[0] > d[1] > d[2]
    assert d[2] < 1e-3


def test_cross_method_agreement(orc):
    field = inputs.kaula_field(10)
    f = orc.Field(field[0], field[1], 10, MU_EARTH, R_EARTH)
    model = orc.make_model(f)
    y0 = inputs.workload("w0", n=4)["y0"]
    out = {}
    for m, rtol in (("dopri5", 1e-13), ("dop853", 1e-13), ("rk87", 1e-13)):
        yf, st, _, _ = orc.propagate_batch(model, orc.make_integ(method=m, rtol=rtol), y0, 0.0, 0.25 * DAY)
        assert np.all(st == 0)
        out[m] = yf
    for m in ("dopri5", "dop853"):
        assert np.max(np.linalg.norm(out[m][:, :3] - out["rk87"][:, :3], axis=1)) < 2e-3


def test_statuses_and_nan_fill(orc):
    f = point_mass_field(orc)
    model = orc.make_model(f, r_min=R_EARTH)
    integ = orc.make_integ(method="dopri5", rtol=1e-10, max_steps=5)
    y0 = inputs.workload("w0", n=3)["y0"]
    yf, st, _, _ = orc.propagate_batch(model, integ, y0, 0.0, DAY)
    assert np.all(st == orc.MAX_STEPS) and np.all(np.isnan(yf))
    # invalid input and below-minimum-radius starts
    bad = y0.copy()
    bad[0, 0] = np.nan
    bad[1, :3] *= 0.5
    integ = orc.make_integ(method="dopri5", rtol=1e-10)
    yf, st, _, _ = orc.propagate_batch(model, integ, bad, 0.0, 600.0)
    assert list(st) == [orc.INVALID_INPUT, orc.BELOW_MIN_RADIUS, 0]
    # a suborbital arc hits the ground during propagation
    sub = np.array([[R_EARTH + 200e3, 0, 0, 0, 3000.0, 0]])
    yf, st, _, _ = orc.propagate_batch(model, integ, sub, 0.0, 3000.0)
    assert st[0] == orc.BELOW_MIN_RADIUS and np.all(np.isnan(yf))
    # max rejects: a tolerance no step can meet
    integ = orc.make_integ(method="dopri5", rtol=0.0, atol_pos=1e-30, atol_vel=1e-30, h_min=1e-3)
    yf, st, _, _ = orc.propagate_batch(model, integ, y0[:1], 0.0, 600.0)
    assert st[0] in (orc.MAX_REJECTS, orc.STEP_UNDERFLOW, orc.MAX_STEPS)


def test_first_step_and_zero_span(orc):
    f = point_mass_field(orc)
    model = orc.make_model(f)
    integ = orc.make_integ(method="rk87", rtol=1e-12)
    y0 = inputs.workload("w0", n=2)["y0"]
    yf, st, steps, _ = orc.propagate_batch(model, integ, y0, 5.0, 5.0)
    assert np.array_equal(yf, y0) and np.all(steps == 0)
    # span <= 60 s: default h0 = span / 2 reaches the end in 2 accepted steps
    yf, st, steps, _ = orc.propagate_batch(model, integ, y0, 0.0, 10.0)
    assert np.all(steps <= 2)


def test_drag_dissipates_energy(orc):
    f = point_mass_field(orc)
    model = orc.make_model(f, omega=0.0, drag="vallado")
    integ = orc.make_integ(method="rk87", rtol=1e-12)
    kep = np.array([[R_EARTH + 300e3, 0.001, 0.5, 0.1, 0.2, 0.3]])
    y0 = kep2eci(kep)
    yf, st, _, _ = orc.propagate_batch(model, integ, y0, 0.0, 0.25 * DAY, bc=np.array([0.02]))
    energy = lambda y: 0.5 * np.sum(y[:, 3:] ** 2, axis=1) - MU_EARTH / np.linalg.norm(y[:, :3], axis=1)  # noqa
    assert st[0] == 0 and energy(yf)[0] < energy(y0)[0]
    # and without drag it is conserved to the integration tolerance
    model0 = orc.make_model(f, omega=0.0, drag="none")
    yf0, _, _, _ = orc.propagate_batch(model0, integ, y0, 0.0, 0.25 * DAY)
    assert abs(energy(yf0)[0] - energy(y0)[0]) < 1e-9 * abs(energy(y0)[0])
    assert energy(y0)[0] - energy(yf)[0] > 1e3 * abs(energy(yf0)[0] - energy(y0)[0])


def test_solve_kepler_identity():
    M = np.linspace(0, 2 * np.pi, 50)
    for e in (0.0, 0.1, 0.7, 0.95):
        E = solve_kepler(M, np.full_like(M, e))
        assert np.max(np.abs(E - e * np.sin(E) - M)) < 1e-14


def test_ground_event_over_the_wgs84_ellipsoid(orc):
    """With a flattening the minimum-radius test is the reference's ground
    event (LY/src/events.rs:87-145): height above the ellipsoid along the
    position's direction.  Over a pole the WGS84 surface is 21.4 km inside
    the equatorial sphere, so a state 110 km above the polar surface clears
    the ellipsoid's 100 km floor but not the sphere R + 100 km; over the
    equator both agree."""
    import bqp_amd

    field = bqp_amd.kaula(4)
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    R, fl = 6378137.0, 1.0 / 298.257223563
    b = R * (1.0 - fl)
    integ = orc.make_integ(method="rk87", rtol=1e-10, canonical=(field.radius, field.mu))

    def status(y0, flat):
        m = orc.make_model(f, r_atm=R, r_min=R + 100e3, ground_flattening=flat)
        _, st, _, _ = orc.propagate_batch(m, integ, np.asarray(y0, dtype=float)[None, :], 0.0, 60.0)
        return int(st[0])

    rp = b + 110e3
    vp = np.sqrt(field.mu / rp)
    polar = [0.0, 0.0, rp, vp, 0.0, 0.0]
    assert status(polar, fl) == 0
    assert status(polar, 0.0) == 6
    re = R + 90e3
    equatorial = [re, 0.0, 0.0, 0.0, np.sqrt(field.mu / re), 0.0]
    assert status(equatorial, fl) == 6 and status(equatorial, 0.0) == 6
    re = R + 110e3
    equatorial = [re, 0.0, 0.0, 0.0, np.sqrt(field.mu / re), 0.0]
    assert status(equatorial, fl) == 0 and status(equatorial, 0.0) == 0
