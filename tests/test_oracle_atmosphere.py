"""Atmosphere and drag of the oracle, pinned independently of every source
file of this project.

* ``orc_density`` (Vallado piecewise model) against a fixture typed from
  Vallado (2013) Table 8-4 (``tests/golden/vallado_2013_table_8_4.json``),
  at every band's base altitude, inside every band and just below its top;
  the fixture itself is checked for continuity at every band edge (the
  published table is continuous there to 1e-4, 1.4e-3 at 25 km), so a
  mistyped digit in it shows as a jump.
* ``orc_density`` (compat model) against the reference's closed form
  (LY/src/rhs.rs:1778-1788: 1.225 exp(-h / 8.5 km), 0 at or above
  1,000 km, 1.225 at or below sea level).
* the drag term of ``orc_rhs`` (LY/src/rhs.rs:1119-1163 compute_drag)
  against -1/2 (Cd A/m) rho |v_rel| v_rel, v_rel = v - w x r, computed here
  in numpy with the fixture's density;
* the drag magnitude along a whole trajectory: in a field rotating
  uniformly about z the Jacobi integral J = v^2/2 - mu/r - w (r x v)_z
  changes only by drag, dJ/dt = a_drag . v_rel = -1/2 bc rho |v_rel|^3, so
  its change over one orbit of the oracle's own sampled integration must
  equal a Simpson quadrature of that rate with the fixture's density.

Planted mistakes (in copies of the oracle source, on the CPU) show these
checks bite: one digit of the table, the factor 1/2 dropped, and v used in
place of v_rel each fail at least one of them.
"""
import json
import os
import shutil

import numpy as np
import pytest
from oracle_variants import build_variant

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OMEGA = 7.292115e-5
RE = 6378137.0
MU = 3.986004418e14


def _fixture():
    with open(os.path.join(HERE, "golden", "vallado_2013_table_8_4.json")) as fh:
        rows = np.array(json.load(fh)["rows"], dtype=np.float64)
    return rows[:, 0], rows[:, 1], rows[:, 2]


def rho_fixture(h_m):
    """Density of the fixture's table at altitude h (m)."""
    h0, r0, H = _fixture()
    hk = np.asarray(h_m, dtype=np.float64) * 1e-3
    b = np.clip(np.searchsorted(h0, hk, side="right") - 1, 0, h0.size - 1)
    out = r0[b] * np.exp(-(hk - h0[b]) / H[b])
    return np.where(hk < 0.0, r0[0], out)


def test_fixture_is_continuous_at_band_edges():
    h0, r0, H = _fixture()
    assert h0.size == 28 and h0[0] == 0.0 and h0[-1] == 1000.0 and np.all(np.diff(h0) > 0)
    top = r0[:-1] * np.exp(-(h0[1:] - h0[:-1]) / H[:-1])
    rel = np.abs(top / r0[1:] - 1.0)
    assert rel[0] < 2e-3, rel[0]          # 25 km edge of the published table
    assert np.all(rel[1:] < 2e-4), rel    # every other edge
    assert np.all(np.diff(r0) < 0) and np.all(H > 0)


def _density_checks(orc):
    """Largest relative difference of orc_density from the fixture over the
    nodes, band interiors, band tops, below sea level and above 1,000 km."""
    h0, r0, H = _fixture()
    hs = list(h0 * 1e3)
    hs += list(0.5 * (h0[:-1] + h0[1:]) * 1e3)
    hs += list(h0[1:] * 1e3 - 1.0)                      # 1 m below each band's top
    hs += [-5000.0, -1.0, 1.2e6, 2.5e6]
    hs = np.array(hs)
    got = np.array([orc.density("vallado", h) for h in hs])
    return float(np.max(np.abs(got / rho_fixture(hs) - 1.0)))


def _drag_term_checks(orc):
    """Largest relative difference of orc_rhs's drag term from the closed
    form, over states at altitudes 155-955 km with prograde, polar and
    retrograde velocities."""
    cbar = np.zeros(6)
    cbar[0] = 1.0
    # GM = 1 m^3/s^2: gravity (~1e-14 m/s^2) does not swamp the drag term in
    # the difference of the two right-hand sides
    f = orc.Field(cbar, np.zeros(6), 2, 1.0, RE)
    with_drag = orc.make_model(f, drag="vallado", omega=OMEGA)
    no_drag = orc.make_model(f, drag="none", omega=OMEGA)
    rng = np.random.default_rng(4)
    worst = 0.0
    for alt in np.linspace(155e3, 955e3, 17):  # off the band edges (the table jumps ~1e-5 there)
        for inc in (0.1, 1.4, 2.9):
            u = rng.uniform(0, 2 * np.pi)
            r = (RE + alt) * np.array([np.cos(u), np.sin(u), 0.0])
            vc = np.sqrt(MU / np.linalg.norm(r))
            tdir = np.array([-np.sin(u), np.cos(u), 0.0])
            v = vc * (np.cos(inc) * tdir + np.sin(inc) * np.array([0.0, 0.0, 1.0]))
            y = np.concatenate([r, v])
            bc = rng.uniform(0.005, 0.02)
            t = rng.uniform(0, 1e4)
            a = orc.rhs(with_drag, bc, t, y)[3:] - orc.rhs(no_drag, bc, t, y)[3:]
            vrel = v - np.cross([0.0, 0.0, OMEGA], r)
            want = -0.5 * bc * rho_fixture(np.linalg.norm(r) - RE) * np.linalg.norm(vrel) * vrel
            worst = max(worst, float(np.linalg.norm(a - want) / np.linalg.norm(want)))
    return worst


def _jacobi(y):
    r = np.linalg.norm(y[..., :3], axis=-1)
    v2 = np.sum(y[..., 3:] ** 2, axis=-1)
    hz = y[..., 0] * y[..., 4] - y[..., 1] * y[..., 3]
    return 0.5 * v2 - MU / r - OMEGA * hz


def _decay_check(orc):
    """Relative difference between the Jacobi-integral change of the
    oracle's drag-only orbit (point-mass field, rotating atmosphere) and the
    quadrature of -1/2 bc rho |v_rel|^3 with the fixture's density.  The
    orbit (perigee 418 km, apogee 432 km) stays inside one band of the
    table, so the integrand is smooth."""
    cbar = np.zeros(6)
    cbar[0] = 1.0
    f = orc.Field(cbar, np.zeros(6), 2, MU, RE)
    model = orc.make_model(f, drag="vallado", omega=OMEGA)
    integ = orc.make_integ(method="rk87", rtol=1e-13)
    a = RE + 425e3
    kep = np.array([a, 0.001, np.radians(30.0), 0.3, 1.1, 0.2])
    y0 = orc.kep2eci(MU, kep)[0]
    period = 2 * np.pi * np.sqrt(a ** 3 / MU)
    m = 400
    te = np.linspace(0.0, period, m + 1)
    bc = 0.02
    out, st, _, _ = orc.propagate_sampled_batch(model, integ, y0[None, :], 0.0, te, np.array([bc]))
    assert st[0] == 0
    y = out[0]
    h = np.linalg.norm(y[:, :3], axis=1) - RE
    assert h.min() > 400e3 and h.max() < 450e3
    vrel = y[:, 3:] - np.cross(np.array([0.0, 0.0, OMEGA]), y[:, :3])
    rate = -0.5 * bc * rho_fixture(h) * np.linalg.norm(vrel, axis=1) ** 3
    dt = te[1] - te[0]
    quad = dt / 3.0 * (rate[0] + rate[-1] + 4.0 * rate[1:-1:2].sum() + 2.0 * rate[2:-1:2].sum())
    dj = _jacobi(y[-1]) - _jacobi(y[0])
    return float(abs(dj / quad - 1.0)), float(dj)


def test_oracle_density_matches_vallado_table(orc):
    assert _density_checks(orc) < 1e-14


def test_oracle_compat_density_closed_form(orc):
    for hk in (-3.0, 0.0, 1e-6, 0.5, 17.0, 120.0, 400.0, 999.999):
        want = 1.225 if hk <= 0.0 else 1.225 * np.exp(-hk / 8.5)
        assert orc.density("compat", hk * 1e3) == pytest.approx(want, rel=1e-15)
    assert orc.density("compat", 1000e3) == 0.0 and orc.density("compat", 2000e3) == 0.0
    assert orc.density("none", 300e3) == 0.0


def test_oracle_drag_term_closed_form(orc):
    assert _drag_term_checks(orc) < 1e-10


def test_oracle_drag_energy_decay_matches_quadrature(orc):
    rel, dj = _decay_check(orc)
    assert dj < 0.0
    assert rel < 1e-6, rel


# ------------------------------------------------------------ planted mistakes

PLANTS = {
    # one digit of the 400-450 km band's nominal density (3.725e-12 -> 3.752e-12)
    "table_digit": ("3.725e-12", "3.752e-12"),
    # the factor 1/2 of the drag acceleration dropped
    "drag_half": ("const double k = -0.5 * bc * rho * vr;", "const double k = -bc * rho * vr;"),
    # the inertial velocity used in place of the velocity relative to the atmosphere
    "v_for_vrel": ("const double vrx = fma(md->omega, y[1], y[3]);\n      const double vry = fma(-md->omega, y[0], "
                   "y[4]);", "const double vrx = y[3];\n      const double vry = y[4];"),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted variants")
@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(orc, tmp_path, name):
    lib = build_variant(name, *PLANTS[name], str(tmp_path))
    with orc.variant(lib):
        dens = _density_checks(orc)
        term = _drag_term_checks(orc)
        decay, _ = _decay_check(orc)
    caught = {"density": dens >= 1e-14, "drag_term": term >= 1e-10, "decay": decay >= 1e-6}
    expected = {"table_digit": ("density", "drag_term", "decay"), "drag_half": ("drag_term", "decay"),
                "v_for_vrel": ("drag_term", "decay")}[name]
    for check in expected:
        assert caught[check], (name, check, dens, term, decay)
