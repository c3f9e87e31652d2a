"""Pins the CPU oracle's gravity (oracle/bqp_oracle.c orc_gravity_bf).

Checks, each chosen so a plausible mistake (dropped term, wrong sign or
index, transposed operand, wrong normalisation) fails at least one:
  * Orekit 13.1.2 Holmes-Featherstone vectors at d/o 5 (reference fixture,
    tolerance 5e-11 m/s^2 + 2e-12 |a| as in LY/src/rhs.rs:892-997);
  * mpmath 60-digit brute force (potential gradient, independent of V/W) at
    N = 8, 20, 50, 70, 120 including the exact pole;
  * point-mass and J2 closed forms;
  * the reference's un-normalised M-G form equals the normalised one;
  * rotation covariance and Laplace's equation (trace of the gradient).
"""
import numpy as np
import pytest

from bqp_amd.inputs import kaula_field, tri_index
from golden_io import mpmath_gravity, orekit

GM = 3.986004418e14
RE = 6378137.0


def test_orekit_noncentral_and_full(orc):
    d = orekit()
    C0 = d["C"].copy()
    C0[0] = 0.0
    full = orc.Field(d["C"], d["S"], d["N"], d["gm"], d["R"])
    nonc = orc.Field(C0, d["S"], d["N"], d["gm"], d["R"])
    for cs in d["cases"]:
        a_full = full.accel(cs["pos"])[0]
        a_nc = nonc.accel(cs["pos"])[0]
        tol_nc = d["atol"] + d["rtol"] * np.linalg.norm(cs["noncentral"])
        tol_full = d["atol"] + d["rtol"] * np.linalg.norm(cs["full"])
        assert np.max(np.abs(a_nc - cs["noncentral"])) <= tol_nc, cs["name"]
        assert np.max(np.abs(a_full - cs["full"])) <= tol_full, cs["name"]
        # much tighter than the reference's own tolerance in practice
        assert np.max(np.abs(a_nc - cs["noncentral"])) <= 1e-14 * np.linalg.norm(cs["full"])


@pytest.mark.parametrize("N", [8, 20, 50, 70, 120])
def test_mpmath_brute_force(orc, N):
    gm, R, cases = mpmath_gravity()
    P, A = cases[N]
    C, S = kaula_field(N)
    f = orc.Field(C, S, N, gm, R)
    a = f.accel(P)
    rel = np.max(np.abs(a - A) / np.linalg.norm(A, axis=1)[:, None])
    assert rel < 1e-14


def test_point_mass_closed_form(orc):
    C = np.zeros(tri_index(9, 0))
    C[0] = 1.0
    f = orc.Field(C, np.zeros_like(C), 8, GM, RE)
    rng = np.random.default_rng(3)
    p = rng.normal(size=(50, 3)) * 7e6
    a = f.accel(p)
    ref = -GM * p / np.linalg.norm(p, axis=1)[:, None] ** 3
    assert np.max(np.abs(a - ref) / np.linalg.norm(ref, axis=1)[:, None]) < 1e-15


def test_j2_closed_form(orc):
    c20 = -4.841653e-4
    C = np.zeros(tri_index(3, 0))
    C[0] = 0.0
    C[tri_index(2, 0)] = c20
    f = orc.Field(C, np.zeros_like(C), 2, GM, RE)
    J2 = -c20 * np.sqrt(5.0)
    rng = np.random.default_rng(4)
    p = rng.normal(size=(50, 3))
    p *= 7.1e6 / np.linalg.norm(p, axis=1)[:, None]
    x, y, z = p.T
    r = np.linalg.norm(p, axis=1)
    k = -1.5 * J2 * GM * RE ** 2 / r ** 5
    zr = 5 * z ** 2 / r ** 2
    ref = np.stack([k * x * (1 - zr), k * y * (1 - zr), k * z * (3 - zr)], axis=1)
    a = f.accel(p)
    assert np.max(np.abs(a - ref)) < 1e-15 * GM / 7.1e6 ** 2


@pytest.mark.parametrize("N", [5, 20, 60])
def test_compat_unnormalised_equals_normalised(orc, N):
    C, S = kaula_field(N)
    fn = orc.Field(C, S, N, GM, RE)
    fc = orc.Field(C, S, N, GM, RE, form="compat")
    rng = np.random.default_rng(N)
    p = rng.normal(size=(64, 3))
    p *= rng.uniform(6.6e6, 8.4e6, 64)[:, None] / np.linalg.norm(p, axis=1)[:, None]
    a, b = fn.accel(p), fc.accel(p)
    assert np.max(np.abs(a - b) / np.linalg.norm(a, axis=1)[:, None]) < 1e-13


def test_rotation_covariance(orc):
    """Rotating the field by angle t about z (C,S of order m rotate by m t) and
    the position by the same angle rotates the acceleration."""
    N = 12
    C, S = kaula_field(N)
    t = 0.731
    C2, S2 = C.copy(), S.copy()
    for n in range(N + 1):
        for m in range(1, n + 1):
            k = tri_index(n, m)
            c, s = np.cos(m * t), np.sin(m * t)
            C2[k] = C[k] * c - S[k] * s
            S2[k] = C[k] * s + S[k] * c
    f1 = orc.Field(C, S, N, GM, RE)
    f2 = orc.Field(C2, S2, N, GM, RE)
    Rz = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    rng = np.random.default_rng(7)
    p = rng.normal(size=(20, 3))
    p *= rng.uniform(6.6e6, 8.4e6, 20)[:, None] / np.linalg.norm(p, axis=1)[:, None]
    a1 = f1.accel(p)
    a2 = f2.accel(p @ Rz.T)
    assert np.max(np.abs(a2 - a1 @ Rz.T)) < 1e-14 * np.max(np.linalg.norm(a1, axis=1))


def test_laplace_trace_zero(orc):
    """Outside the masses the potential is harmonic: div(a) = 0 (central FD)."""
    N = 20
    C, S = kaula_field(N)
    f = orc.Field(C, S, N, GM, RE)
    p0 = np.array([4.1e6, -3.3e6, 4.4e6])
    h = 2.0
    tr = 0.0
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        tr += (f.accel(p0 + e)[0][k] - f.accel(p0 - e)[0][k]) / (2 * h)
    scale = 2 * GM / np.linalg.norm(p0) ** 3  # size of one diagonal entry
    assert abs(tr) < 1e-7 * scale


def test_low_order_field_truncation(orc):
    """mmax < nmax drops exactly the orders above mmax."""
    N = 10
    C, S = kaula_field(N)
    Ct, St = C.copy(), S.copy()
    for n in range(N + 1):
        for m in range(4, n + 1):
            Ct[tri_index(n, m)] = 0.0
            St[tri_index(n, m)] = 0.0
    p = np.array([[5.0e6, 4.0e6, 2.0e6]])
    a1 = orc.Field(C, S, N, GM, RE, mmax=3).accel(p)
    a2 = orc.Field(Ct, St, N, GM, RE).accel(p)
    assert np.max(np.abs(a1 - a2)) < 1e-16 * np.linalg.norm(a1)
