"""Gravity field containers and loaders (host side).

``load_icgem`` reads ICGEM ``.gfc`` files (``gfc n m C S [sigmas]`` lines with
``earth_gravity_constant`` / ``radius`` / ``norm`` header keys) and the plain
``n m C S`` text the reference bundles; both are fully normalised.  This
replaces the reference's coefficient loader
(dust_transfer/crates/lightyear_odeint_rs/src/config.rs:126-223), which
de-normalises on load; here coefficients stay normalised because the
recursion is the normalised one.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .inputs import kaula_field, tri_index

HERE = os.path.dirname(os.path.abspath(__file__))
GOCE_D15 = os.path.join(HERE, "data", "goce_dir_r6_d15.txt")
GOCE_GM = 3.986004415e14
GOCE_R = 6378136.46


@dataclass
class Field:
    """Normalised coefficients, triangular idx n(n+1)/2 + m, n = 0..nmax."""

    cbar: np.ndarray
    sbar: np.ndarray
    nmax: int
    mu: float
    radius: float
    mmax: int | None = None

    def __post_init__(self):
        self.cbar = np.ascontiguousarray(self.cbar, dtype=np.float64)
        self.sbar = np.ascontiguousarray(self.sbar, dtype=np.float64)
        if self.mmax is None:
            self.mmax = self.nmax
        need = tri_index(self.nmax + 1, 0)
        if self.cbar.size < need or self.sbar.size < need:
            raise ValueError("coefficient arrays shorter than the triangular size")

    def truncated(self, nmax: int, mmax: int | None = None) -> "Field":
        if nmax > self.nmax:
            raise ValueError(f"field has degree {self.nmax} < {nmax}")
        k = tri_index(nmax + 1, 0)
        return Field(self.cbar[:k].copy(), self.sbar[:k].copy(), nmax, self.mu, self.radius,
                     min(nmax, self.mmax if mmax is None else mmax))


def kaula(nmax: int, seed: int = 0, mu: float = 3.986004418e14, radius: float = 6378137.0) -> Field:
    C, S = kaula_field(nmax, seed)
    return Field(C, S, nmax, mu, radius)


def load_icgem(path: str, nmax: int | None = None, mu: float | None = None,
               radius: float | None = None) -> Field:
    """Parse an ICGEM .gfc (or plain 'n m C S') file into a normalised Field."""
    def num(tok: str) -> float:  # Fortran exponents (1.0D-05) are common in ICGEM files
        return float(tok.replace("D", "E").replace("d", "e"))

    hdr_mu, hdr_r, norm = None, None, "fully_normalized"
    rows = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            key = parts[0].lower()
            if key == "earth_gravity_constant":
                hdr_mu = num(parts[1])
            elif key == "radius":
                hdr_r = num(parts[1])
            elif key == "norm":
                norm = parts[1].lower()
            elif key in ("gfc", "gfct"):
                rows.append((int(parts[1]), int(parts[2]), num(parts[3]), num(parts[4])))
            elif key.lstrip("-").isdigit() and len(parts) >= 4:
                rows.append((int(parts[0]), int(parts[1]), num(parts[2]), num(parts[3])))
    if norm not in ("fully_normalized", "fully_normalised"):
        raise ValueError(f"unsupported normalisation {norm!r}")
    if not rows:
        raise ValueError(f"no coefficients in {path}")
    top = max(r[0] for r in rows)
    nmax = top if nmax is None else int(nmax)
    if nmax > top:
        raise ValueError(f"file holds degree {top} < requested {nmax}")
    C = np.zeros(tri_index(nmax + 1, 0))
    S = np.zeros_like(C)
    for n, m, c, s in rows:
        if n <= nmax and 0 <= m <= n:
            C[tri_index(n, m)] = c
            S[tri_index(n, m)] = s if m > 0 else 0.0
    C[0] = 1.0
    return Field(C, S, nmax, mu if mu is not None else (hdr_mu or GOCE_GM),
                 radius if radius is not None else (hdr_r or GOCE_R))


def goce(nmax: int) -> Field:
    """ESA GOCE GO_CONS_GCF_2_DIR_R6 truncated to degree 15 (CC BY 4.0), the
    field the reference's Python binding bakes in."""
    return load_icgem(GOCE_D15, nmax=nmax, mu=GOCE_GM, radius=GOCE_R)
