"""Order conditions of every Butcher tableau (the published coefficient sets).

For each method the propagated weights satisfy all rooted-tree conditions up
to the method's order and fail at order+1, the embedded weights likewise at
their order, and the rows satisfy c_i = sum_j a_ij.  The same check runs on
the copy compiled into libbqp.so (bqp_get_tableau) when it is built.
"""
import os

import numpy as np
import pytest

from rk_trees import residuals, trees

CASES = [("dopri5", 5, 4), ("dop853", 8, 5), ("rk87", 8, 7)]


def test_tree_counts():
    assert [len(trees(p)) for p in range(1, 9)] == [1, 1, 2, 4, 9, 20, 48, 115]


@pytest.mark.parametrize("name,p,pe", CASES)
def test_order_conditions(orc, name, p, pe):
    t = orc.tableau(name)
    A, b, e = t["a"], t["b"], t["e"]
    assert t["order"] == p and t["order_err"] == pe
    assert np.max(np.abs(A.sum(axis=1) - t["c"])) < 5e-15
    assert np.allclose(np.triu(A), 0.0)
    r = residuals(A, b, p + 1)
    assert max(r[:p]) < 5e-15, r
    assert r[p] > 1e-8  # not order p+1
    rh = residuals(A, b - e, pe + 1)
    assert max(rh[:pe]) < 5e-15, rh
    assert rh[pe] > 1e-8
    if t["has_e3"]:
        r3 = residuals(A, b - t["e3"], 4)
        assert max(r3[:3]) < 5e-15 and r3[3] > 1e-8


def test_fsal_structure(orc):
    t = orc.tableau("dopri5")
    assert t["fsal"]
    assert np.array_equal(t["a"][-1, :], t["b"]) or np.allclose(t["a"][-1, :6], t["b"][:6], rtol=0, atol=0)
    assert t["c"][-1] == 1.0


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "bqp_amd", "libbqp.so")),
                    reason="libbqp.so not built")
@pytest.mark.parametrize("name", ["dopri5", "dop853", "rk87"])
def test_library_tableau_matches_oracle(orc, name):
    import bqp_amd

    a = bqp_amd.tableau(name)
    o = orc.tableau(name)
    for k in ("c", "a", "b", "e", "e3"):
        assert np.array_equal(a[k], o[k]), k
