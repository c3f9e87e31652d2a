"""Step-controller behaviours of the oracle, restating the reference's unit
tests LY/src/odesolve/basic_tests.rs:394-500 on its own test system y' = y
(ExpSystem; exact solution y0 e^t).  The reference runs them with Vern9,
which this project does not carry, so its exact step counts (47 and 45) do
not transfer; each test instead checks the behaviour by the observation
that distinguishes it from the alternative, from the oracle's per-attempt
trace (t, h, error norm, threshold, accepted, next h):

* solver preset (solver.rs:445-529): a rejected step does not write the PI
  memory, so the step after a reject is sized with the error of the last
  ACCEPTED step; the compat preset (lightyear_compat.rs:559-575) does write
  it, so there the reject's error sizes it;
* h_min force-accept: with h pinned at h_min = h_max and an unsatisfiable
  tolerance every attempt fails the error test and is accepted anyway;
* saturation: with a loose tolerance every step is clipped by h_max, with a
  tight one the controller sizes the steps below the cap;
* the growth and shrink clamps ([0.2, min(1 + 20/q, 5)], at most 2 right
  after a reject; [0.1, 0.5] on a reject).
"""
import shutil

import numpy as np
import pytest
from oracle_variants import build_variant

MU, RE = 3.986004418e14, 6378137.0


@pytest.fixture(scope="module")
def exp_model(orc):
    cbar = np.zeros(6)
    cbar[0] = 1.0
    f = orc.Field(cbar, np.zeros(6), 2, MU, RE)
    return orc.make_model(f, formulation="test_exp")


def _pi_candidates(tr, k, eps, q, compat, h_min, h_max):
    """Next step after accepted attempt k (which follows a reject) under the
    two hypotheses: PI memory = last accepted error, or = the reject's."""
    acc = tr[:, 4]
    p = max(j for j in range(k - 1) if acc[j] == 1.0)
    hs, en = tr[k, 1], tr[k, 2]
    out = []
    for prev in (tr[p, 2], tr[k - 1, 2]):
        if compat:
            fac = 0.9 * (eps / en) ** (0.7 / 5.0) * (prev / eps) ** (0.4 / 5.0)
            grow = 2.0
        else:
            fac = 0.9 * (eps / en) ** (0.7 / q) * (prev / eps) ** (0.4 / q)
            grow = min(min(1.0 + 20.0 / q, 5.0), 2.0)
        out.append(min(max(hs * min(max(fac, 0.2), grow), h_min), h_max))
    return out


def _accepts_after_rejects(tr):
    acc = tr[:, 4]
    return [k for k in range(1, len(tr)) if acc[k] == 1.0 and acc[k - 1] == 0.0 and acc[:k - 1].any()]


def test_solver_preset_rejects_do_not_seed_pi_memory(orc, exp_model):
    """Absolute eps 1e-14, h0 = h_max = 2 s on [0, 4] (the reference's
    fixture), RK8(7)13M: the step after the mid-run reject follows the last
    accepted error, and the endpoint meets the closed form."""
    integ = orc.make_integ(method="rk87", control="absolute", eps=1e-14, h0=2.0, h_max=2.0)
    yf, st, tr = orc.propagate_traced(exp_model, integ, np.ones(6), 0.0, 4.0)
    assert st.status == 0 and st.rejects > 0
    ks = _accepts_after_rejects(tr)
    assert ks, "fixture no longer has an accept after a reject with earlier accepts"
    q = orc.tableau("rk87")["order_err"] + 1
    for k in ks:
        from_accepted, from_reject = _pi_candidates(tr, k, 1e-14, q, False, 1e-12, 2.0)
        assert abs(from_accepted / from_reject - 1.0) > 1e-3  # the fixture discriminates
        assert tr[k, 5] == pytest.approx(from_accepted, rel=1e-13)
    assert np.all(np.abs(yf - np.exp(4.0)) < 1e-13 * np.exp(4.0))


def test_compat_preset_rejects_do_seed_pi_memory(orc, exp_model):
    """The DOPRI5 compat loop writes err_prev on a reject
    (lightyear_compat.rs:559-575): the step after a reject follows the
    reject's error."""
    for tf in (4.0, 8.0):
        integ = orc.make_integ(method="dopri5", control="compat", eps=1e-12)
        yf, st, tr = orc.propagate_traced(exp_model, integ, np.ones(6), 0.0, tf)
        assert st.status == 0
        ks = _accepts_after_rejects(tr)
        assert ks
        for k in ks:
            from_accepted, from_reject = _pi_candidates(tr, k, 1e-12, 5.0, True, 1e-10, tf)
            assert abs(from_accepted / from_reject - 1.0) > 1e-3
            assert tr[k, 5] == pytest.approx(from_reject, rel=1e-13)
        assert abs(yf[0] / np.exp(tf) - 1.0) < 1e-12


@pytest.mark.parametrize("method", ["rk87", "dop853", "dopri5"])
def test_steps_at_h_min_are_force_accepted(orc, exp_model, method):
    """h0 = h_min = h_max = 1 and eps = 1e-30 on [0, 10]: each of the 10
    attempts fails the error test (err > eps) and is accepted because h
    cannot shrink; none is rejected, and the run reports success."""
    integ = orc.make_integ(method=method, control="absolute", eps=1e-30, h0=1.0, h_min=1.0, h_max=1.0)
    yf, st, tr = orc.propagate_traced(exp_model, integ, np.ones(6), 0.0, 10.0)
    assert st.status == 0 and st.steps == 10 and st.rejects == 0
    assert len(tr) == 10 and np.all(tr[:, 4] == 1.0)
    assert np.all(tr[:, 2] > tr[:, 3])       # every one failed the error test
    assert np.all(tr[:, 1] == 1.0)
    assert np.allclose(tr[:, 0], np.arange(10.0))


def test_saturation_separates_the_cap_from_the_controller(orc, exp_model):
    """Loose tolerance under a tight cap: every accepted step is the cap.
    Tight tolerance under a loose cap: the controller sizes the steps below
    it.  Both directions rule out a cap that never binds or always binds."""
    capped = orc.make_integ(method="rk87", control="absolute", eps=1.0, h0=0.25, h_max=0.25)
    _, st, tr = orc.propagate_traced(exp_model, capped, np.ones(6), 0.0, 4.0)
    assert st.status == 0 and st.steps == 16 and st.rejects == 0
    assert np.all(tr[:, 1] == 0.25) and np.all(tr[:, 5] == 0.25)
    free = orc.make_integ(method="rk87", control="absolute", eps=1e-14, h0=0.25, h_max=4.0)
    _, st, tr = orc.propagate_traced(exp_model, free, np.ones(6), 0.0, 4.0)
    acc = tr[tr[:, 4] == 1.0]
    assert st.status == 0 and np.all(acc[:, 5] < 4.0)
    assert np.mean(acc[:-1, 1] < 0.5) > 0.9


@pytest.mark.parametrize("method,control,eps", [("rk87", "absolute", 1e-14), ("dop853", "absolute", 1e-13),
                                                 ("dopri5", "compat", 1e-12), ("rk87", "scaled", 1e-13)])
def test_step_ratio_clamps(orc, exp_model, method, control, eps):
    """Accepted: next/current in [0.2, min(1 + 20/q, 5)], at most 2 right
    after a reject (compat: 5, and 2 after a reject); rejected: in
    [0.1, 0.5] (h_max never binds here)."""
    integ = orc.make_integ(method=method, control=control, eps=eps, rtol=eps, h0=3.0, h_max=100.0)
    _, st, tr = orc.propagate_traced(exp_model, integ, np.ones(6), 0.0, 6.0)
    assert st.status == 0 and st.rejects > 0
    q = 5.0 if control == "compat" else orc.tableau(method)["order_err"] + 1.0
    grow = 5.0 if control == "compat" else min(1.0 + 20.0 / q, 5.0)
    ratio = tr[:, 5] / tr[:, 1]
    acc = tr[:, 4] == 1.0
    after_rej = np.zeros(len(tr), dtype=bool)
    after_rej[1:] = ~acc[:-1]
    last = np.zeros(len(tr), dtype=bool)
    last[-1] = True  # the last step lands on tf, its next h is not used
    a = acc & ~last
    assert np.all(ratio[a] >= 0.2 - 1e-15) and np.all(ratio[a] <= grow + 1e-15)
    assert np.all(ratio[a & after_rej] <= 2.0 + 1e-15)
    assert np.all(ratio[~acc] >= 0.1 - 1e-15) and np.all(ratio[~acc] <= 0.5 + 1e-15)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted variant")
def test_planted_pi_memory_write_on_reject_fails(orc, tmp_path):
    """A copy of the oracle whose solver preset writes the PI memory on a
    reject (as only the compat preset should) fails the solver-preset test."""
    lib = build_variant("pi_seed", "      if (compat) {\n        err_prev = en;\n        have_prev = 1;\n      }",
                        "      {\n        err_prev = en;\n        have_prev = 1;\n      }", str(tmp_path))
    with orc.variant(lib):
        cbar = np.zeros(6)
        cbar[0] = 1.0
        f = orc.Field(cbar, np.zeros(6), 2, MU, RE)
        model = orc.make_model(f, formulation="test_exp")
        integ = orc.make_integ(method="rk87", control="absolute", eps=1e-14, h0=2.0, h_max=2.0)
        _, st, tr = orc.propagate_traced(model, integ, np.ones(6), 0.0, 4.0)
    q = orc.tableau("rk87")["order_err"] + 1
    ks = _accepts_after_rejects(tr)
    assert ks
    mism = [k for k in ks if abs(tr[k, 5] / _pi_candidates(tr, k, 1e-14, q, False, 1e-12, 2.0)[0] - 1.0) > 1e-13]
    assert mism, "the planted PI-memory write went unnoticed"
