"""The compiled-in second-order tableau tables (bqp_amd/csrc/second_order_tables.h).

For y = (r, v) the kernels store only stage accelerations and use
R_s = r + h c^_s v + h^2 sum_j (A.A)_sj A_j and the matching (b.A), (e.A),
(e3.A) weights.  Every entry of the generated header is checked against the
same products formed in exact rational arithmetic from the published
coefficients (the generator sums in long double: within one ulp of the
result plus the long-double rounding bound 16 * 2^-64 * sum|terms|, which
matters where the terms cancel), the packed
stage rows against their definition, and the committed header against a
fresh run of its generator.  Planted mistakes (a transposed product, b.A
taken as A.b) are shown to fail.
"""
import os
import re
import shutil
import subprocess
import tempfile
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "bqp_amd", "csrc")
HEADER = os.path.join(CSRC, "second_order_tables.h")
METHODS = ["dopri5", "dop853", "rk87"]  # the order of the three entries in the header
MAXS = 13
CHUNK = 4
CHUNKS = (MAXS - 1 + CHUNK - 1) // CHUNK


def header_values():
    """The header's initializer as three flat lists of doubles."""
    text = open(HEADER).read()
    body = text[text.index("#define BQP_SECOND_ORDER_TABLES"):]
    nums = [float.fromhex(t) for t in re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+", body)]
    per = MAXS + MAXS * MAXS + 3 * MAXS + 3 + MAXS * CHUNKS * 2 * CHUNK
    assert len(nums) == 3 * per
    out = []
    for m in range(3):
        v = nums[m * per:(m + 1) * per]
        k = 0

        def take(n):
            nonlocal k
            r = v[k:k + n]
            k += n
            return r
        d = {"csum": take(MAXS), "a2": np.array(take(MAXS * MAXS)).reshape(MAXS, MAXS), "ba": take(MAXS),
             "ea": take(MAXS), "e3a": take(MAXS)}
        d["bsum"], d["esum"], d["e3sum"] = take(3)
        d["sc"] = np.array(take(MAXS * CHUNKS * 2 * CHUNK)).reshape(MAXS, CHUNKS, 2 * CHUNK)
        out.append(d)
    return out


def exact(t):
    S = t["stages"]
    A = [[Fraction(float(t["a"][i, j])) for j in range(S)] for i in range(S)]
    b = [Fraction(float(x)) for x in t["b"][:S]]
    e = [Fraction(float(x)) for x in t["e"][:S]]
    e3 = [Fraction(float(x)) for x in t["e3"][:S]]

    def dot(terms):
        # exact value and the sum of magnitudes that bounds long-double rounding
        return sum(terms), sum(abs(x) for x in terms)
    return {"S": S,
            "a2": [[dot([A[i][k] * A[k][j] for k in range(S)]) for j in range(S)] for i in range(S)],
            "csum": [dot(r) for r in A],
            "ba": [dot([b[j] * A[j][i] for j in range(S)]) for i in range(S)],
            "ea": [dot([e[j] * A[j][i] for j in range(S)]) for i in range(S)],
            "e3a": [dot([e3[j] * A[j][i] for j in range(S)]) for i in range(S)],
            "bsum": dot(b), "esum": dot(e), "e3sum": dot(e3)}


def close(got, want_mag):
    """got within one ulp of the exact value plus the long-double summation
    bound; a zero whose terms are all zero must be exactly zero."""
    want, mag = want_mag
    if mag == 0:
        return got == 0.0
    err = abs(Fraction(got) - want)
    return err <= Fraction(float(np.spacing(abs(float(want))))) + Fraction(16, 2 ** 64) * mag


@pytest.mark.parametrize("m", range(3))
def test_header_matches_exact_products(orc, m):
    h = header_values()[m]
    x = exact(orc.tableau(METHODS[m]))
    S = x["S"]
    for i in range(S):
        assert close(h["csum"][i], x["csum"][i]), (i, h["csum"][i])
        for key in ("ba", "ea", "e3a"):
            assert close(h[key][i], x[key][i]), (key, i)
        for j in range(S):
            assert close(h["a2"][i, j], x["a2"][i][j]), (i, j)
    for key in ("bsum", "esum", "e3sum"):
        assert close(h[key], x[key]), key
    # rows of unused stages are zero
    assert not np.any(h["a2"][S:]) and not np.any(h["csum"][S:])


@pytest.mark.parametrize("m", range(3))
def test_packed_stage_rows(orc, m):
    h = header_values()[m]
    t = orc.tableau(METHODS[m])
    S = t["stages"]
    for s in range(MAXS):
        for j in range(MAXS - 1):
            k, q = divmod(j, CHUNK)
            a = t["a"][s, j] if (s < S and j < s) else 0.0
            a2 = h["a2"][s, j] if (s < S and j < s) else 0.0
            assert h["sc"][s, k, 2 * q] == a and h["sc"][s, k, 2 * q + 1] == a2, (s, j)


def test_planted_mistakes_fail(orc):
    """A transposed stage product or b.A formed as A.b differ from the header."""
    h = header_values()[2]
    t = orc.tableau("rk87")
    S = t["stages"]
    A = t["a"][:S, :S]
    assert not np.allclose((A.T @ A), h["a2"][:S, :S], rtol=1e-12, atol=1e-15)
    assert not np.allclose(A @ t["b"][:S], h["ba"][:S], rtol=1e-12, atol=1e-15)
    # the same tolerance as the exact comparison rejects them
    x = exact(dict(t, a=t["a"].T.copy()))
    assert not all(close(h["a2"][i, j], x["a2"][i][j]) for i in range(S) for j in range(S))
    assert not all(close(h["ba"][i], x["ba"][i]) for i in range(S))


@pytest.mark.skipif(shutil.which(os.environ.get("CXX", "g++")) is None, reason="no host C++ compiler")
def test_committed_header_is_current():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        out = os.path.join(d, "tables.h")
        subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-o", exe,
                        os.path.join(CSRC, "gen_second_order.cpp")], check=True)
        subprocess.run([exe, out], check=True)
        assert open(out).read() == open(HEADER).read()
