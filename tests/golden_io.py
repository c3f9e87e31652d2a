"""Readers for the committed golden fixtures (hex binary64 strings)."""
import json
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hexd(s: str) -> float:
    return struct.unpack(">d", bytes.fromhex(s[2:].rjust(16, "0")))[0]


def orekit():
    with open(os.path.join(GOLDEN, "orekit_dir_r6_d5.json")) as fh:
        d = json.load(fh)
    gm, R = hexd(d["gm_m3_s2"]), hexd(d["reference_radius_m"])
    N = int(d["runtime_degree"])
    C = np.zeros((N + 1) * (N + 2) // 2)
    S = np.zeros_like(C)
    for n, m, c, s in d["coefficients_nmcs"]:
        if n <= N:
            C[n * (n + 1) // 2 + m] = c
            S[n * (n + 1) // 2 + m] = s
    C[0] = 1.0
    cases = []
    for cs in d["cases"]:
        cases.append({
            "name": cs["name"],
            "pos": np.array([hexd(x) for x in cs["position_m"]]),
            "noncentral": np.array([hexd(x) for x in cs["orekit_noncentral_m_s2"]]),
            "point_mass": np.array([hexd(x) for x in cs["point_mass_m_s2"]]),
            "full": np.array([hexd(x) for x in cs["full_m_s2"]]),
        })
    return dict(gm=gm, R=R, N=N, C=C, S=S, cases=cases, atol=hexd(d["absolute_tolerance_m_s2"]),
                rtol=hexd(d["relative_tolerance"]))


def mpmath_gravity():
    with open(os.path.join(GOLDEN, "mpmath_gravity.json")) as fh:
        d = json.load(fh)
    out = {}
    for N, c in d["cases"].items():
        P = np.array([[hexd(x) for x in p] for p in c["positions"]])
        A = np.array([[hexd(x) for x in a] for a in c["acc"]])
        out[int(N)] = (P, A)
    return hexd(d["gm"]), hexd(d["radius"]), out
