"""CPU checks of the Encke/Battin formulation of the oracle: the two-body
baseline against the Kepler closed form, Battin's correction against the
direct difference in extended precision, and whole Encke propagations
against Cowell propagations of the same problem."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs

MU = 3.986004418e14


def _elliptic(rng, n):
    a = rng.uniform(6.6e6, 4.2e7, n)
    e = rng.uniform(0.0, 0.7, n)
    kep = np.column_stack([a, e, np.deg2rad(rng.uniform(0, 180, n)), rng.uniform(0, 2 * np.pi, n),
                           rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)])
    return kep


def test_two_body_matches_kepler_closed_form(orc):
    rng = np.random.default_rng(0)
    kep = _elliptic(rng, 60)
    s0 = orc.kep2eci(MU, kep)
    for k in range(60):
        n = np.sqrt(MU / kep[k, 0] ** 3)
        for dt in (0.0, 17.0, 3600.0, -5400.0, 7.3 * 2 * np.pi / n):
            later = kep[k].copy()
            later[5] = kep[k, 5] + n * dt
            want = orc.kep2eci(MU, later[None, :])[0]
            got, _ = orc.two_body(MU, s0[k], dt)
            assert np.linalg.norm(got[:3] - want[:3]) <= 1e-9 * np.linalg.norm(want[:3]) + 1e-6, (k, dt)
            assert np.linalg.norm(got[3:] - want[3:]) <= 1e-9 * np.linalg.norm(want[3:]) + 1e-9


def test_two_body_conserves_energy_and_momentum(orc):
    rng = np.random.default_rng(1)
    s0 = orc.kep2eci(MU, _elliptic(rng, 20))
    for s in s0:
        out, x = orc.two_body(MU, s, 12345.0)
        en = lambda v: 0.5 * np.dot(v[3:], v[3:]) - MU / np.linalg.norm(v[:3])
        assert abs(en(out) - en(s)) <= 1e-12 * abs(en(s))
        assert np.allclose(np.cross(out[:3], out[3:]), np.cross(s[:3], s[3:]), rtol=1e-12)
        # a warm start at the solution returns the same state
        again, x2 = orc.two_body(MU, s, 12345.0, x=x)
        assert np.allclose(again, out, rtol=1e-14, atol=1e-9) and abs(x2 - x) < 1e-12


def test_two_body_rejects_non_ellipse(orc):
    r = 7.0e6
    s = np.array([r, 0, 0, 0, np.sqrt(2 * MU / r) * 1.01, 0])  # hyperbolic
    with pytest.raises(ValueError):
        orc.two_body(MU, s, 100.0)


def test_battin_matches_extended_precision_difference(orc):
    rng = np.random.default_rng(2)
    ld = np.longdouble
    for _ in range(200):
        d = rng.normal(size=3)
        rb = d / np.linalg.norm(d) * rng.uniform(6.6e6, 8.4e6)
        dr = rng.normal(size=3) * 10 ** rng.uniform(0, 4)
        acc = orc.battin(MU, rb, dr)
        r = rb.astype(ld) + dr.astype(ld)
        rbl = rb.astype(ld)
        want = ld(MU) * (rbl / (rbl @ rbl) ** ld(1.5) - r / (r @ r) ** ld(1.5))
        assert np.max(np.abs(acc - want.astype(float))) <= 1e-12 * float(np.max(np.abs(want))) + 1e-22


def _both(orc, field, method, drag, span, n, encke_eps, cowell_rtol, control="absolute"):
    w = inputs.workload("w1", n=n, seed=3)
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    cow = orc.make_model(f, drag=drag)
    enc = orc.make_model(f, drag=drag, formulation="encke")
    ic = orc.make_integ(method=method, rtol=cowell_rtol, canonical=(field.radius, field.mu))
    ie = orc.make_integ(method=method, control=control, eps=encke_eps, canonical=(field.radius, field.mu))
    bc = w["bc"][:n] if drag != "none" else None
    yc, sc, _, _ = orc.propagate_batch(cow, ic, w["y0"][:n], 0.0, span, bc)
    ye, se, steps, evals = orc.propagate_batch(enc, ie, w["y0"][:n], 0.0, span, bc)
    return yc, sc, ye, se, steps


def test_encke_matches_cowell(orc):
    yc, sc, ye, se, steps = _both(orc, bqp_amd.kaula(8), "rk87", "none", 86400.0, 8, 1e-7, 1e-13)
    assert np.all(sc == 0) and np.all(se == 0)
    err = np.linalg.norm(ye[:, :3] - yc[:, :3], axis=1)
    assert err.max() < 1e-3, err  # metres after one day (16 rebased segments)
    assert np.all(steps >= 16)


def test_encke_dopri5_compat_preset_matches_cowell(orc):
    yc, sc, ye, se, _ = _both(orc, bqp_amd.kaula(8), "dopri5", "none", 10800.0, 4, 1e-9, 1e-12,
                              control="compat")
    assert np.all(se == 0)
    assert np.linalg.norm(ye[:, :3] - yc[:, :3], axis=1).max() < 0.05


def test_encke_detects_a_planted_battin_sign_error(orc):
    # the Encke/Cowell agreement above fails when the correction is negated
    w = inputs.workload("w1", n=1, seed=3)
    field = bqp_amd.kaula(4)
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    y0 = w["y0"][0]
    base, _ = orc.two_body(MU, y0, 600.0)
    dr = np.array([300.0, -200.0, 100.0])
    good = orc.battin(field.mu, base[:3], dr)
    direct = field.mu * (base[:3] / np.linalg.norm(base[:3]) ** 3 - (base[:3] + dr) / np.linalg.norm(base[:3] + dr) ** 3)
    assert np.allclose(good, direct, rtol=1e-9)
    assert not np.allclose(-good, direct, rtol=1e-3)


def test_step_size_carry_across_rebases(orc):
    """A rebased segment opens with the step size the controller asked for at
    the end of the previous one (LY/src/integrator.rs:341-402).  The carry
    changes the step sequence (results differ from cold restarts, by less
    than the Encke tolerance moves them from Cowell) at about the same cost
    here (the span/100 opening guess is already near the controller's step),
    and there is no carry where the reference has none: the DOPRI5 compat
    loop and step caps of 60 s or less."""
    field = bqp_amd.kaula(8)
    w = inputs.workload("w1", n=8, seed=3)
    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    enc = orc.make_model(f, formulation="encke")
    cow = orc.make_model(f)
    can = (field.radius, field.mu)

    def run(model, **kw):
        integ = orc.make_integ(canonical=can, **kw)
        y, s, steps, evals = orc.propagate_batch(model, integ, w["y0"], 0.0, 86400.0)
        assert np.all(s == 0)
        return y, evals

    yc, _ = run(cow, method="rk87", rtol=1e-13)
    y1, e1 = run(enc, method="rk87", control="absolute", eps=1e-7)
    y0, e0 = run(enc, method="rk87", control="absolute", eps=1e-7, hcarry=False)
    assert not np.array_equal(e1, e0) and not np.array_equal(y1, y0)
    assert abs(e1.sum() / e0.sum() - 1.0) < 0.02, (e1, e0)
    assert np.linalg.norm(y1[:, :3] - yc[:, :3], axis=1).max() < 1e-3
    assert np.linalg.norm(y1[:, :3] - y0[:, :3], axis=1).max() < 1e-3
    # no carry: a 60 s cap, or the compat loop -> identical to the switch turned off
    for kw in ({"method": "rk87", "control": "absolute", "eps": 1e-7, "h_max": 60.0},
               {"method": "dopri5", "control": "compat", "eps": 1e-9}):
        ya, ea = run(enc, **kw)
        yb, eb = run(enc, hcarry=False, **kw)
        assert np.array_equal(ya, yb) and np.array_equal(ea, eb), kw
