"""Independent FP64 gravitational potential for invariant checks (numpy).

U(r) = (GM/r) sum_n (R/r)^n sum_m Pbar_nm(sin phi) (C cos m lam + S sin m lam),
with Pbar from the standard normalised degree recursion (not the V/W
Cunningham form), evaluated in the body frame rotated by theta about z.
"""
import numpy as np

from bqp_amd.elements import MU_EARTH, R_EARTH
from bqp_amd.inputs import tri_index


def pbar(N, s):
    c = np.sqrt(max(0.0, 1.0 - s * s))
    P = np.zeros((N + 1, N + 1))
    P[0, 0] = 1.0
    for m in range(1, N + 1):
        f = np.sqrt(3.0) if m == 1 else np.sqrt((2.0 * m + 1.0) / (2.0 * m))
        P[m, m] = f * c * P[m - 1, m - 1]
    for m in range(0, N + 1):
        if m + 1 <= N:
            P[m + 1, m] = np.sqrt(2.0 * m + 3.0) * s * P[m, m]
        for n in range(m + 2, N + 1):
            a = np.sqrt((2.0 * n + 1) * (2.0 * n - 1) / ((n - m) * (n + m)))
            b = np.sqrt((2.0 * n + 1) * (n + m - 1) * (n - m - 1) / ((n - m) * (n + m) * (2.0 * n - 3)))
            P[n, m] = a * s * P[n - 1, m] - b * P[n - 2, m]
    return P


def potential_bf(field, N, p, mu=MU_EARTH, radius=R_EARTH):
    C, S = field
    x, y, z = p
    r = np.sqrt(x * x + y * y + z * z)
    P = pbar(N, z / r)
    lam = np.arctan2(y, x)
    tot = 0.0
    q = radius / r
    for n in range(N + 1):
        row = 0.0
        for m in range(n + 1):
            row += P[n, m] * (C[tri_index(n, m)] * np.cos(m * lam) + S[tri_index(n, m)] * np.sin(m * lam))
        tot += q ** n * row
    return mu / r * tot


def potential_rotating(field, N, r_inertial, theta):
    c, s = np.cos(theta), np.sin(theta)
    x, y, z = r_inertial
    return potential_bf(field, N, (c * x + s * y, -s * x + c * y, z))


def potential_rotating_batch(field, N, r_inertial, theta, mu=MU_EARTH, radius=R_EARTH):
    """potential_rotating for many positions at once (rows of r_inertial),
    sweeping one order m at a time so memory stays O(rows)."""
    C, S = field
    r_inertial = np.asarray(r_inertial, dtype=np.float64)
    c, s = np.cos(theta), np.sin(theta)
    x = c * r_inertial[:, 0] + s * r_inertial[:, 1]
    y = -s * r_inertial[:, 0] + c * r_inertial[:, 1]
    z = r_inertial[:, 2]
    r = np.sqrt(x * x + y * y + z * z)
    sphi = z / r
    cphi = np.sqrt(np.maximum(0.0, 1.0 - sphi * sphi))
    lam = np.arctan2(y, x)
    q = radius / r
    tot = np.zeros_like(r)
    pmm = np.ones_like(r)  # Pbar_mm
    for m in range(N + 1):
        if m >= 1:
            f = np.sqrt(3.0) if m == 1 else np.sqrt((2.0 * m + 1.0) / (2.0 * m))
            pmm = f * cphi * pmm
        cm, sm = np.cos(m * lam), np.sin(m * lam)
        p2 = np.zeros_like(r)
        p1 = pmm
        qn = q ** m
        col = qn * p1 * (C[tri_index(m, m)] * cm + S[tri_index(m, m)] * sm)
        for n in range(m + 1, N + 1):
            if n == m + 1:
                p = np.sqrt(2.0 * m + 3.0) * sphi * p1
            else:
                a = np.sqrt((2.0 * n + 1) * (2.0 * n - 1) / ((n - m) * (n + m)))
                b = np.sqrt((2.0 * n + 1) * (n + m - 1) * (n - m - 1) / ((n - m) * (n + m) * (2.0 * n - 3)))
                p = a * sphi * p1 - b * p2
            qn = qn * q
            col = col + qn * p * (C[tri_index(n, m)] * cm + S[tri_index(n, m)] * sm)
            p2, p1 = p1, p
        tot = tot + col
    return mu / r * tot
