"""Seeded synthetic inputs for the propagation workloads (frozen draw order).

Every generator uses ``numpy.random.Generator(PCG64(seed))`` and draws in the
order written here; changing the order changes every fixture, so the order is
part of the interface.

Workloads (trajectory counts and durations are the full sizes; tests slice
prefixes):
  w0  1,024 LEO orbits, 1 day, DOPRI5(4), WRMS rtol = atol = 1e-10 in canonical
      units, degree/order 8 Kaula field, no drag.
  w1  65,536 LEO orbits, 7 days, RK8(7)13M, rtol 1e-12, degree/order 70,
      Vallado piecewise exponential drag, Cd*A/m ~ U(0.005, 0.02) m^2/kg.
  w2  262,144 mixed LEO/MEO/GEO, log-uniform 1 h..30 d, RK8(7)13M rtol 1e-11,
      degree/order 20, gravity only.
  w3  as w1 with the learned residual surrogate (degree 120 training field).
  w4  131,072 = 32 x 4,096 population candidates, arcs U(0.5, 10) d,
      degree/order 50 + drag, delta-v U(-50, 50) m/s per axis.
"""
from __future__ import annotations

import numpy as np

from .elements import MU_EARTH, R_EARTH, kep2eci

C20 = -4.841653e-4
DAY = 86400.0


def tri_index(n, m):
    return n * (n + 1) // 2 + m


def kaula_field(nmax: int, seed: int = 0, c20: float = C20):
    """Normalised Cbar/Sbar (triangular, idx n(n+1)/2+m) from the Kaula rule.

    C00 = 1, degree 1 zero, C20 fixed; every other Cnm/Snm for n = 2..nmax is
    N(0, (1e-5/n^2)^2).  Draw order: degree by degree, each degree one vector
    of 2n+1 normals laid out [C_n0, C_n1, S_n1, C_n2, S_n2, ...]; the C20 slot
    is drawn and replaced by ``c20``.  The field of degree N is therefore a
    prefix of every higher-degree field from the same seed.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    size = tri_index(nmax + 1, 0)
    C = np.zeros(size)
    S = np.zeros(size)
    C[0] = 1.0
    for n in range(2, nmax + 1):
        z = rng.normal(0.0, 1e-5 / (n * n), 2 * n + 1)
        C[tri_index(n, 0)] = z[0]
        for m in range(1, n + 1):
            C[tri_index(n, m)] = z[2 * m - 1]
            S[tri_index(n, m)] = z[2 * m]
    C[tri_index(2, 0)] = c20
    return C, S


def leo_elements(n: int, seed: int = 0, alt=(400e3, 1200e3), ecc=(0.0, 0.02), inc_deg=(0.0, 98.0),
                 rng=None):
    """[a, e, i, raan, argp, M] (m, rad); draw order alt, e, i, raan, argp, M."""
    rng = rng if rng is not None else np.random.Generator(np.random.PCG64(seed))
    h = rng.uniform(alt[0], alt[1], n)
    e = rng.uniform(ecc[0], ecc[1], n)
    i = np.deg2rad(rng.uniform(inc_deg[0], inc_deg[1], n))
    raan = np.deg2rad(rng.uniform(0.0, 360.0, n))
    argp = np.deg2rad(rng.uniform(0.0, 360.0, n))
    M = np.deg2rad(rng.uniform(0.0, 360.0, n))
    return np.stack([R_EARTH + h, e, i, raan, argp, M], axis=-1)


def canonical_atol(rtol: float, mu: float = MU_EARTH, radius: float = R_EARTH):
    """atol equal to rtol in canonical units: DU = R, VU = sqrt(mu/R)."""
    return rtol * radius, rtol * np.sqrt(mu / radius)


def workload(name: str, n: int | None = None, seed: int = 0) -> dict:
    """Inputs of a named workload: states (n,6), t0 (n,), tf (n,), bc (n,), field, solver."""
    if name == "w0":
        n = 1024 if n is None else n
        kep = leo_elements(n, seed)
        y0 = kep2eci(kep)
        return dict(name=name, y0=y0, t0=np.zeros(n), tf=np.full(n, DAY), bc=np.zeros(n), nmax=8,
                    drag="none", method="dopri5", rtol=1e-10)
    if name in ("w1", "w3"):
        n = 65536 if n is None else n
        rng = np.random.Generator(np.random.PCG64(seed))
        kep = leo_elements(n, rng=rng)
        bc = rng.uniform(0.005, 0.02, n)
        y0 = kep2eci(kep)
        return dict(name=name, y0=y0, t0=np.zeros(n), tf=np.full(n, 7 * DAY), bc=bc, nmax=70,
                    drag="vallado", method="rk87", rtol=1e-12)
    if name == "w2":
        n = 262144 if n is None else n
        rng = np.random.Generator(np.random.PCG64(seed))
        cls = rng.uniform(0.0, 1.0, n)
        is_leo = cls < 0.60
        is_meo = (cls >= 0.60) & (cls < 0.85)
        a = np.where(is_leo, R_EARTH + rng.uniform(300e3, 2000e3, n),
                     np.where(is_meo, rng.uniform(20000e3, 30000e3, n), 42164e3))
        e = np.where(is_leo | is_meo, rng.uniform(0.0, 0.02, n), rng.uniform(0.0, 1e-3, n))
        inc = np.deg2rad(np.where(is_leo | is_meo, rng.uniform(0.0, 98.0, n), rng.uniform(0.0, 1.0, n)))
        raan = np.deg2rad(rng.uniform(0.0, 360.0, n))
        argp = np.deg2rad(rng.uniform(0.0, 360.0, n))
        M = np.deg2rad(rng.uniform(0.0, 360.0, n))
        dur = np.exp(rng.uniform(np.log(3600.0), np.log(30 * DAY), n))
        y0 = kep2eci(np.stack([a, e, inc, raan, argp, M], axis=-1))
        return dict(name=name, y0=y0, t0=np.zeros(n), tf=dur, bc=np.zeros(n), nmax=20, drag="none",
                    method="rk87", rtol=1e-11, orbit_class=np.where(is_leo, 0, np.where(is_meo, 1, 2)))
    if name == "w4":
        n = 131072 if n is None else n
        rng = np.random.Generator(np.random.PCG64(seed))
        kep = leo_elements(n, rng=rng)
        y0 = kep2eci(kep)
        y0[:, 3:] += rng.uniform(-50.0, 50.0, (n, 3))
        dur = rng.uniform(0.5 * DAY, 10 * DAY, n)
        bc = rng.uniform(0.005, 0.02, n)
        return dict(name=name, y0=y0, t0=np.zeros(n), tf=dur, bc=bc, nmax=50, drag="vallado",
                    method="rk87", rtol=1e-12, generations=32, population=4096)
    raise ValueError(f"unknown workload {name!r}")
