"""Size-independent checks at the full 65,536-row batch shape.

Gravity only, the field rotating uniformly about z: every trajectory keeps its
Jacobi integral J = |v|^2/2 - U(r_bf) - w (x v_y - y v_x).  U comes from an
independent numpy potential (standard normalised Legendre degree recursion,
not the kernel's rescaled V/W form), so a wrong coefficient, sign, index or
frame rotation in the GPU acceleration breaks J by orders of magnitude more
than the integrator's own drift (at most 2.6e-13 relative over a day for the
oracle at rtol 1e-12).  A strided sample of the same rows is also checked
against the oracle."""
import numpy as np
import pytest

import bqp_amd
from bqp_amd import inputs
from gpu_util import announce, oracle_pair, pos_err
from potential_ref import potential_rotating_batch

pytestmark = pytest.mark.gpu

OMEGA = 7.292115e-5
DAY = 86400.0


def jacobi(field, N, y, theta):
    return (0.5 * (y[:, 3:] ** 2).sum(axis=1) - potential_rotating_batch(field, N, y[:, :3], theta)
            - OMEGA * (y[:, 0] * y[:, 4] - y[:, 1] * y[:, 3]))


def test_full_batch_jacobi_integral_and_oracle_sample(orc):
    announce("65,536 LEO x 1 day, N=70, gravity only: Jacobi integral of every row, 16 rows vs oracle")
    n = 65536
    y0 = inputs.workload("w1", n=n)["y0"]
    cbar, sbar = inputs.kaula_field(70)
    field = bqp_amd.kaula(70)
    assert np.array_equal(field.cbar, cbar) and np.array_equal(field.sbar, sbar)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, omega=OMEGA, theta0=0.0) as p:
        p.propagate_batch_final(y0, 0.0, np.full(n, DAY))  # trains the hand-out order
        yf, st, info = p.propagate_batch_final(y0, 0.0, np.full(n, DAY))
    assert np.all(st == 0)
    j0 = jacobi((cbar, sbar), 70, y0, 0.0)
    j1 = jacobi((cbar, sbar), 70, yf, OMEGA * DAY)
    rel = np.abs(j1 - j0) / np.abs(j0)
    announce(f"Jacobi drift: max {rel.max():.3e}, median {np.median(rel):.3e}")
    assert rel.max() < 1e-11
    idx = np.arange(0, n, 4096)
    _, model, integ = oracle_pair(orc, field, rtol=1e-12)
    yo, so, _, _ = orc.propagate_batch(model, integ, y0[idx], 0.0, DAY)
    assert np.all(so == 0)
    assert pos_err(yf[idx], yo).max() < 1e-2


def test_full_batch_drag_lowers_every_jacobi_integral():
    """With drag, dJ/dt = v_rel . a_drag = -(1/2)(Cd A/m) rho |v_rel|^3 <= 0:
    J must fall for every row, and by more for the lower, draggier orbits."""
    announce("65,536 LEO x 1 day, N=70, Vallado drag: every Jacobi integral falls")
    n = 65536
    w = inputs.workload("w1", n=n)
    cbar, sbar = inputs.kaula_field(70)
    field = bqp_amd.kaula(70)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado", omega=OMEGA, theta0=0.0) as p:
        yf, st, _ = p.propagate_batch_final(w["y0"], 0.0, np.full(n, DAY), w["bc"])
    assert np.all(st == 0)
    drop = jacobi((cbar, sbar), 70, w["y0"], 0.0) - jacobi((cbar, sbar), 70, yf, OMEGA * DAY)
    announce(f"J drop: min {drop.min():.3e}, median {np.median(drop):.3e}, max {drop.max():.3e} m^2/s^2")
    assert np.all(drop > 0.0)
    # per unit ballistic coefficient, the lowest tenth of the orbits lose more than the highest tenth
    a = -field.mu / (2.0 * two_body_energy(w["y0"], field.mu))
    k = drop / w["bc"]
    lo, hi = a < np.quantile(a, 0.1), a > np.quantile(a, 0.9)
    assert np.median(k[lo]) > 10.0 * np.median(k[hi])


def two_body_energy(y, mu):
    """Two-body specific energy (for ranking orbits by size)."""
    return 0.5 * (y[:, 3:] ** 2).sum(axis=1) - mu / np.linalg.norm(y[:, :3], axis=1)


def test_full_batch_encke_jacobi_integral_and_oracle_sample(orc):
    """The Encke formulation (deviation from a two-body baseline, rebased every
    5,400 s) at the full batch shape: 6 hours, so every row crosses four
    rebases.  Gravity only, so each row keeps its Jacobi integral (the
    oracle's Encke drift over the same arcs is at most 5.6e-14), and a strided
    sample matches the oracle's Encke to 1 cm."""
    announce("65,536 LEO x 6 h, N=70, Encke, absolute eps 1e-6: Jacobi integral of every row, 16 rows vs oracle")
    n = 65536
    span = 0.25 * DAY
    y0 = inputs.workload("w1", n=n)["y0"]
    cbar, sbar = inputs.kaula_field(70)
    field = bqp_amd.kaula(70)
    with bqp_amd.Propagator(field, method="rk87", control="absolute", eps=1e-6, formulation="encke",
                            omega=OMEGA, theta0=0.0) as p:
        yf, st, _ = p.propagate_batch_final(y0, 0.0, np.full(n, span))
    assert np.all(st == 0)
    j0 = jacobi((cbar, sbar), 70, y0, 0.0)
    j1 = jacobi((cbar, sbar), 70, yf, OMEGA * span)
    rel = np.abs(j1 - j0) / np.abs(j0)
    announce(f"Encke Jacobi drift: max {rel.max():.3e}, median {np.median(rel):.3e}")
    assert rel.max() < 1e-11
    idx = np.arange(0, n, 4096)
    f = orc.Field(cbar, sbar, 70, field.mu, field.radius)
    model = orc.make_model(f, omega=OMEGA, formulation="encke")
    integ = orc.make_integ("rk87", control="absolute", eps=1e-6, canonical=(field.radius, field.mu))
    yo, so, _, _ = orc.propagate_batch(model, integ, y0[idx], 0.0, span)
    assert np.all(so == 0)
    assert pos_err(yf[idx], yo).max() < 1e-2
