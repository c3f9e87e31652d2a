"""Brute-force high-precision gravity golden vectors (mpmath).

Independent of the Cunningham V/W formulation the oracle and the kernels use:

  U(r) = (GM/r) * sum_{n<=N} (R/r)^n sum_{m<=n} Pbar_nm(sin phi)
                                     * (Cbar_nm cos m lam + Sbar_nm sin m lam)

is evaluated at 60 significant digits, with Pbar_nm built from the
un-normalised Ferrers functions (no Condon-Shortley phase) by the textbook
degree recursion and the exact factorial normalisation
N_nm = sqrt((2-d_m0)(2n+1)(n-m)!/(n+m)!).  That recursion is itself checked
against mpmath.legenp (hypergeometric evaluation) for n <= 20 before use.
The acceleration is grad U by fourth-order central differences in Cartesian
coordinates (step 1e-18 r), so there is no pole singularity.

Fields: the seeded Kaula field (bqp_amd.inputs.kaula_field, seed 0) at
N = 8, 20, 50, 70, 120; GM = 3.986004418e14, R = 6378137.
Positions: seeded radii U(6578, 8378) km with uniform directions, plus the
exact north pole and an equatorial point.

Output: mpmath_gravity.json (hex binary64 of the rounded results).
Run: python tests/golden/gen_mpmath_gravity.py   (a few minutes)
"""
import json
import os
import struct
import sys

import mpmath as mp
import numpy as np

sys.path.insert(0, os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "..")))
from bqp_amd.inputs import kaula_field, tri_index  # noqa: E402

mp.mp.dps = 60
GM = mp.mpf("3.986004418e14")
RE = mp.mpf(6378137)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mpmath_gravity.json")


def hexd(x: float) -> str:
    return "0x" + struct.pack(">d", float(x)).hex()


def pbar_table(N, s):
    """Fully normalised Pbar_nm(s), n,m <= N, at mp precision."""
    c = mp.sqrt(1 - s * s)
    P = {}
    for m in range(N + 1):
        # un-normalised P_mm = (2m-1)!! c^m (no Condon-Shortley phase)
        pmm = mp.mpf(1)
        for k in range(1, m + 1):
            pmm *= (2 * k - 1) * c
        P[(m, m)] = pmm
        if m + 1 <= N:
            P[(m + 1, m)] = s * (2 * m + 1) * pmm
        for n in range(m + 2, N + 1):
            P[(n, m)] = ((2 * n - 1) * s * P[(n - 1, m)] - (n + m - 1) * P[(n - 2, m)]) / (n - m)
    out = {}
    for (n, m), v in P.items():
        nf = mp.sqrt((1 if m == 0 else 2) * (2 * n + 1) * mp.factorial(n - m) / mp.factorial(n + m))
        out[(n, m)] = nf * v
    return out


def check_against_legenp(N=20):
    s = mp.mpf("0.37")
    tab = pbar_table(N, s)
    worst = mp.mpf(0)
    for n in range(N + 1):
        for m in range(n + 1):
            ref = (-1) ** m * mp.legenp(n, m, s, type=2)
            nf = mp.sqrt((1 if m == 0 else 2) * (2 * n + 1) * mp.factorial(n - m) / mp.factorial(n + m))
            worst = max(worst, abs(tab[(n, m)] - nf * ref) / max(mp.mpf(1e-30), abs(nf * ref)))
    assert worst < mp.mpf("1e-40"), worst
    return worst


def potential(N, C, S, x, y, z):
    r = mp.sqrt(x * x + y * y + z * z)
    s = z / r
    lam = mp.atan2(y, x)
    P = pbar_table(N, s)
    tot = mp.mpf(0)
    q = RE / r
    qn = mp.mpf(1)
    for n in range(N + 1):
        row = mp.mpf(0)
        for m in range(n + 1):
            cnm = mp.mpf(C[tri_index(n, m)])
            snm = mp.mpf(S[tri_index(n, m)])
            if cnm == 0 and snm == 0:
                continue
            row += P[(n, m)] * (cnm * mp.cos(m * lam) + snm * mp.sin(m * lam))
        tot += qn * row
        qn *= q
    return GM / r * tot


def accel(N, C, S, p):
    x, y, z = (mp.mpf(float(v)) for v in p)
    r = mp.sqrt(x * x + y * y + z * z)
    h = r * mp.mpf("1e-18")
    g = []
    for k in range(3):
        def U(d):
            q = [x, y, z]
            q[k] += d
            return potential(N, C, S, *q)
        g.append((8 * (U(h) - U(-h)) - (U(2 * h) - U(-2 * h))) / (12 * h))
    return g


def positions(count, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(6578e3, 8378e3, count)
    v = rng.normal(size=(count, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pts = (v * r[:, None]).tolist()
    pts.append([0.0, 0.0, 7000e3])          # exact pole
    pts.append([6878137.0, 0.0, 0.0])       # equator
    return pts


def main():
    print("legenp check worst rel:", check_against_legenp())
    out = {"schema": "bqp_mpmath_gravity_v1", "gm": hexd(3.986004418e14), "radius": hexd(6378137.0),
           "field": "kaula seed 0 (bqp_amd.inputs.kaula_field)", "cases": {}}
    for N, count in ((8, 6), (20, 6), (50, 4), (70, 4), (120, 2)):
        C, S = kaula_field(N, seed=0)
        pts = positions(count, seed=1000 + N)
        accs = []
        for p in pts:
            a = accel(N, C, S, p)
            accs.append([hexd(float(v)) for v in a])
            print(N, p, [float(v) for v in a], flush=True)
        out["cases"][str(N)] = {"positions": [[hexd(v) for v in p] for p in pts], "acc": accs}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
