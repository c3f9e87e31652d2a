"""Pins the oracle's sampled mode (orc_propagate_sampled_batch), which
restates the reference's t_eval batches (LY/src/batch.rs:247-390) and
`propagate(times_s=...)` (DT/src/lib.rs:237-353): steps are shortened onto
each sample time and the controller carries through the samples.

* Kepler closed form at every sample of a two-body arc (Cowell and Encke):
  a sample written at the wrong index or time, or a state taken before the
  landing step, fails.
* With the same opening step (h0 set: by default it is a fraction of the
  whole span), the arc up to the first sample is a final-state arc to that
  time, so the first sample equals orc_propagate_batch to t_1 bit for bit;
  a one-point grid equals the final-state call.
* A sample at t0 returns the initial state; samples beyond a failed arc's
  end are NaN.
* Samples of a harmonic field with Vallado drag agree with final-state calls
  to each sample time within the integration tolerance.
"""
import os
import shutil
import sys

import numpy as np
import pytest

from bqp_amd import inputs
from bqp_amd.elements import MU_EARTH, R_EARTH, kep2eci

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle_variants import build_variant  # noqa: E402

DAY = 86400.0


def point_mass_field(orc, N=2):
    C = np.zeros((N + 1) * (N + 2) // 2)
    C[0] = 1.0
    return orc.Field(C, np.zeros_like(C), N, MU_EARTH, R_EARTH)


def kepler_state(kep, t):
    k = kep.copy()
    k[5] = k[5] + np.sqrt(MU_EARTH / k[0] ** 3) * t
    return kep2eci(k)


def _kepler_errors(orc, formulation, control, eps):
    """Largest position and velocity differences from the Kepler closed form
    over every sample of six two-body LEO arcs."""
    f = point_mass_field(orc)
    model = orc.make_model(f, omega=0.0, formulation=formulation)
    integ = orc.make_integ(method="rk87", control=control, rtol=1e-12, eps=eps)
    kep = inputs.leo_elements(6, seed=21)
    y0 = kep2eci(kep)
    te = np.array([0.0, 17.0, 600.0, 3000.0, 3001.5, 9000.0, 0.4 * DAY])
    ys, st, _, _ = orc.propagate_sampled_batch(model, integ, y0, 0.0, te)
    assert np.all(st == 0)
    want = np.stack([np.stack([kepler_state(kep[i], t) for t in te]) for i in range(6)])
    with np.errstate(over="ignore", invalid="ignore"):  # a planted mistake leaves slots unwritten
        dr = np.nan_to_num(np.linalg.norm(ys[..., :3] - want[..., :3], axis=-1), nan=np.inf)
        dv = np.nan_to_num(np.linalg.norm(ys[..., 3:] - want[..., 3:], axis=-1), nan=np.inf)
    return dr.max(), dv.max(), np.array_equal(ys[:, 0], y0)


@pytest.mark.parametrize("formulation,control,eps", [("cowell", "scaled", 1e-8), ("encke", "absolute", 1e-7)])
def test_every_sample_matches_kepler(orc, formulation, control, eps):
    dr, dv, first = _kepler_errors(orc, formulation, control, eps)
    assert first
    assert dr < 2e-3 and dv < 2e-6, (formulation, dr, dv)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted variant")
def test_planted_sample_index_shift_fails(orc, tmp_path):
    """A sample stored one slot late (the last slot keeps its own) must fail
    the closed-form check."""
    lib = build_variant("sample_shift", "memcpy(stops->out + 6 * stops->next, tot, sizeof(tot));",
                        "memcpy(stops->out + 6 * (stops->next + 1 < stops->m ? stops->next + 1 : stops->next), "
                        "tot, sizeof(tot));", str(tmp_path))
    with orc.variant(lib):
        dr, dv, _ = _kepler_errors(orc, "cowell", "scaled", 1e-8)
    assert dr > 1.0 and dv > 1e-3


def test_first_sample_is_the_final_state_arc(orc):
    import bqp_amd
    fld = bqp_amd.kaula(12)
    f = orc.Field(fld.cbar, fld.sbar, fld.nmax, fld.mu, fld.radius)
    model = orc.make_model(f)
    integ = orc.make_integ(method="dop853", rtol=1e-11, h0=30.0)
    w = inputs.workload("w1", n=8)
    te = np.array([2500.0, 7000.0, 20000.0])
    ys, st, _, _ = orc.propagate_sampled_batch(model, integ, w["y0"], 0.0, te)
    y1, s1, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, te[0])
    assert np.all(st == 0) and np.all(s1 == 0)
    assert np.array_equal(ys[:, 0], y1)
    # a one-point grid is the final-state call
    yo, so, steps_o, _ = orc.propagate_sampled_batch(model, integ, w["y0"], 0.0, te[2:])
    yf, sf, steps_f, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, te[2])
    assert np.array_equal(yo[:, 0], yf) and np.array_equal(steps_o, steps_f)
    # later samples: the controller carried through the forced landings, so
    # they differ from fresh final-state arcs only at the tolerance level
    for m in (1, 2):
        ym, _, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, te[m])
        assert np.max(np.linalg.norm(ys[:, m, :3] - ym[:, :3], axis=1)) < 1e-3


def test_drag_samples_agree_with_final_state_calls(orc):
    import bqp_amd
    fld = bqp_amd.kaula(20)
    f = orc.Field(fld.cbar, fld.sbar, fld.nmax, fld.mu, fld.radius)
    model = orc.make_model(f, drag="vallado")
    integ = orc.make_integ(method="rk87", rtol=1e-12)
    w = inputs.workload("w1", n=6, seed=4)
    te = np.linspace(0.0, 0.25 * DAY, 7)[1:]
    ys, st, _, _ = orc.propagate_sampled_batch(model, integ, w["y0"], 0.0, te, w["bc"])
    assert np.all(st == 0)
    for m, t in enumerate(te):
        yf, sf, _, _ = orc.propagate_batch(model, integ, w["y0"], 0.0, t, w["bc"])
        assert np.all(sf == 0)
        assert np.max(np.linalg.norm(ys[:, m, :3] - yf[:, :3], axis=1)) < 1e-3, m


def test_samples_after_a_failed_arc_are_nan(orc):
    f = point_mass_field(orc)
    model = orc.make_model(f, omega=0.0)
    integ = orc.make_integ(method="rk87", rtol=1e-12, max_steps=40)
    kep = inputs.leo_elements(3, seed=5)
    y0 = kep2eci(kep)
    te = np.array([10.0, 0.05 * DAY, 2.0 * DAY])
    ys, st, _, _ = orc.propagate_sampled_batch(model, integ, y0, 0.0, te)
    assert np.all(st != 0)
    assert np.all(np.isfinite(ys[:, 0]))  # reached before the step budget ran out
    assert np.all(np.isnan(ys[:, 2]))
