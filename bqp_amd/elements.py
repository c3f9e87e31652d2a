"""Keplerian <-> inertial conversions (numpy, vectorised, FP64).

Host-side input preparation for the propagator.  Mirrors the reference's
``kep_to_eci`` / ``eci_to_kep`` bindings (dust_transfer/src/lib.rs:200-221,
backed by satpy_core ``kep2eci_impl`` SC/src/lib.rs:1616-1736) in meaning:
elements ``[a, e, i, raan, argp, anomaly]``.  The SI functions here take
metres and radians; the reference-shaped wrappers in ``bqp_amd`` take km and
degrees.
"""
from __future__ import annotations

import numpy as np

MU_EARTH = 3.986004418e14  # m^3/s^2
R_EARTH = 6378137.0  # m
OMEGA_EARTH = 7.292115e-5  # rad/s


def solve_kepler(M, e, tol=1e-15, maxiter=60):
    """Eccentric anomaly from mean anomaly by Newton iteration (elliptic)."""
    M = np.asarray(M, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    E = np.where(e < 0.8, M, np.pi * np.ones_like(M))
    for _ in range(maxiter):
        d = (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
        E = E - d
        if np.all(np.abs(d) < tol):
            break
    return E


def kep2eci(kep, mu=MU_EARTH, anomaly="mean"):
    """Elements (..., 6) = [a (m), e, i, raan, argp, anomaly] (rad) -> state (..., 6) m, m/s."""
    kep = np.asarray(kep, dtype=np.float64)
    a, e, inc, raan, argp, anom = (kep[..., k] for k in range(6))
    if anomaly == "mean":
        E = solve_kepler(anom, e)
    elif anomaly == "true":
        E = 2.0 * np.arctan2(np.sqrt(1.0 - e) * np.sin(anom / 2.0),
                             np.sqrt(1.0 + e) * np.cos(anom / 2.0))
    else:
        raise ValueError("anomaly must be 'mean' or 'true'")
    cE, sE = np.cos(E), np.sin(E)
    sq = np.sqrt(1.0 - e * e)
    xp = a * (cE - e)
    yp = a * sq * sE
    nmean = np.sqrt(mu / (a * a * a))
    den = 1.0 - e * cE
    vxp = -a * nmean * sE / den
    vyp = a * nmean * sq * cE / den
    cO, sO, ci, si, cw, sw = np.cos(raan), np.sin(raan), np.cos(inc), np.sin(inc), np.cos(argp), np.sin(argp)
    p = np.stack([cO * cw - sO * sw * ci, sO * cw + cO * sw * ci, sw * si], axis=-1)
    q = np.stack([-cO * sw - sO * cw * ci, -sO * sw + cO * cw * ci, cw * si], axis=-1)
    r = p * xp[..., None] + q * yp[..., None]
    v = p * vxp[..., None] + q * vyp[..., None]
    return np.concatenate([r, v], axis=-1)


def eci2kep(state, mu=MU_EARTH, anomaly="true"):
    """State (..., 6) -> elements [a, e, i, raan, argp, anomaly] (m, rad)."""
    s = np.asarray(state, dtype=np.float64)
    r, v = s[..., :3], s[..., 3:]
    rn = np.linalg.norm(r, axis=-1)
    vn2 = np.sum(v * v, axis=-1)
    h = np.cross(r, v)
    hn = np.linalg.norm(h, axis=-1)
    k = np.zeros_like(h)
    k[..., 2] = 1.0
    nvec = np.cross(k, h)
    nn = np.linalg.norm(nvec, axis=-1)
    evec = ((vn2 - mu / rn)[..., None] * r - np.sum(r * v, axis=-1)[..., None] * v) / mu
    e = np.linalg.norm(evec, axis=-1)
    energy = vn2 / 2.0 - mu / rn
    a = -mu / (2.0 * energy)
    inc = np.arccos(np.clip(h[..., 2] / hn, -1.0, 1.0))
    raan = np.where(nn > 0, np.arctan2(nvec[..., 1], nvec[..., 0]), 0.0) % (2 * np.pi)
    with np.errstate(invalid="ignore", divide="ignore"):
        argp = np.where(
            (nn > 0) & (e > 0),
            np.arctan2(np.sum(np.cross(nvec, evec) * h, axis=-1) / hn, np.sum(nvec * evec, axis=-1)),
            0.0,
        ) % (2 * np.pi)
        nu = np.where(
            e > 0,
            np.arctan2(np.sum(np.cross(evec, r) * h, axis=-1) / hn, np.sum(evec * r, axis=-1)),
            0.0,
        ) % (2 * np.pi)
    if anomaly == "true":
        anom = nu
    else:
        E = 2.0 * np.arctan2(np.sqrt(1.0 - e) * np.sin(nu / 2.0), np.sqrt(1.0 + e) * np.cos(nu / 2.0))
        anom = (E - e * np.sin(E)) % (2 * np.pi)
    return np.stack([a, e, inc, raan, argp, anom], axis=-1)
