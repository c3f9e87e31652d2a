"""Throughput benchmark: orbit-days/s of the batched propagator on one MI355X.

Workload (one "step" = one pass of the hot path over the whole batch):
  65,536 LEO orbits (altitude U(400,1200) km, e U(0,0.02), i U(0,98) deg,
  angles U(0,360) deg, seed 0), 7 days each (458,752 orbit-days), RK8(7)13M
  WRMS rtol 1e-12 (atol = rtol in canonical units), degree/order 70 seeded
  Kaula field, Vallado piecewise exponential drag with Cd*A/m U(0.005,0.02).
Inputs are resident in HBM before the timed region; the timed region is the
device call (cost-sort + persistent propagation kernel + row statistics).
The untimed warm-up calls also teach the context's cost model (evaluations
per revolution by altitude and inclination) that orders the rows of the
timed calls; a call made before any warm-up orders them by revolutions.

`value` is the plain FP64 form: W untimed warm-up passes, then K timed
passes.  Its JSON line (with roofline and cpu_baseline) is printed as soon as
it is measured.  The split-precision form, the learned surrogate (trained on
the device first, untimed; BF16 network, exact fallback on a failed sampled
check) and the Encke formulation (the reference's own: deviation from a
two-body baseline rebased every 5,400 s, absolute error control) then run on
the same inputs, one timed pass each after a one-hour warm-up pass, as long
as their projected time fits --budget-s; a second line repeats the headline
with their rates and final-position differences from FP64 under "forms".
The CPU baseline is the project's FP64 C restatement (oracle/) on a bounded
sample of the same workload, run before any GPU work.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N]
       [--days D] [--forms fp64,split,surrogate,encke] [--no-cpu]
       [--budget-s S] [--workload w1|w2|w3|w4|floors]
       (w2: ragged batch; w3: surrogate study; w4: NSGA-II population;
        floors: error floors of the surrogate and of the split form)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "orbit-days/s, 65,536 LEO @ deg-70: plain FP64 vs surrogate @ ≤1 m err"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak


def stage_flops(method: str) -> float:
    """Algorithmic FLOPs per step attempt outside the RHS: stage sums
    (2 per fma, 1 per h*a), solution + error sums (Kahan: 4 per term),
    error norm and controller (~40)."""
    import bqp_amd

    t = bqp_amd.tableau(method)
    a = t["a"]
    nz = int(np.count_nonzero(a))
    f = nz * (1 + 6 * 2)
    f += int(np.count_nonzero(t["b"])) * (1 + 6 * 2)
    f += int(np.count_nonzero(t["e"])) * (1 + 6 * 5)
    f += int(np.count_nonzero(t["e3"])) * (1 + 6 * 5)
    return float(f + 40)


def cpu_baseline(w, field, threads: int, target_s: float, days: float = 1.0) -> dict:
    """The oracle (FP64 C restatement, OpenMP over rows) on the first rows of
    the same workload for `days` each.  A pilot of 2 rows per thread sizes
    the timed sample to about `target_s` seconds; the rate is the timed
    sample's orbit-days over its wall time."""
    from oracle import oracle as orc

    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    model = orc.make_model(f, drag="vallado")
    integ = orc.make_integ(method="rk87", rtol=1e-12, canonical=(field.radius, field.mu))
    n_all = w["y0"].shape[0]

    def timed(rows):
        t0 = time.perf_counter()
        _, st, _, _ = orc.propagate_batch(model, integ, w["y0"][:rows], 0.0, days * 86400.0, w["bc"][:rows],
                                          threads=threads)
        return time.perf_counter() - t0, int(np.sum(st == 0))

    pilot = min(n_all, 2 * threads)
    pw, _ = timed(pilot)
    # whole rounds of `threads` rows (one row per thread at a time)
    rows = int(min(n_all, max(pilot, round(pilot * target_s / max(pw, 1e-3) / threads) * threads)))
    wall, ok = timed(rows)
    return {
        "value": rows * days / wall,
        "unit": "orbit-days/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {rows} orbits of the same workload for {days:g} day(s) each ({ok}/{rows} ok), FP64 C "
                  f"restatement of the same algorithm (oracle/), OpenMP, one row per thread at a time, wall "
                  f"{wall:.1f} s (sized by a {pilot}-row pilot of {pw:.1f} s)",
    }


MLP_FLOPS = 2.0 * (16 * 256 + 3 * 256 * 256 + 256 * 3)  # per network evaluation
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
FP8_PEAK_TFLOPS = 5000.0   # MI355X dense FP8 MFMA


def train_surrogate(prop, rows: int, steps: int, batch: int = 8192, seed: int = 0):
    """Set up and train the propagator's surrogate on `rows` seeded device
    samples of its own field (residual above degree 4, |r| in [6578, 8378]
    km).  Returns (surrogate, features, targets, report)."""
    import bqp_amd
    from bqp_amd.device import synchronize

    sur = bqp_amd.Surrogate(prop, nlow=4, r_lo=6578.0e3, r_hi=8378.0e3, seed=seed)
    t = time.perf_counter()
    feat, targ, _ = sur.dataset(rows, seed=seed + 1)
    synchronize()
    t_data = time.perf_counter() - t
    t = time.perf_counter()
    losses = sur.fit(feat, targ, steps=steps, batch=batch, lr_max=2e-3, lr_min=1e-5, chunk=500, seed=seed)
    synchronize()
    t_train = time.perf_counter() - t
    rep = {"samples": rows, "data_s": t_data, "steps": steps, "batch": batch, "train_s": t_train,
           "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-50:]))}
    print(f"[bench] surrogate trained: {rows} samples in {t_data:.2f} s, {steps} x {batch} steps in "
          f"{t_train:.1f} s, loss {rep['loss_first']:.3g} -> {rep['loss_last']:.3g}", file=sys.stderr, flush=True)
    return sur, feat, targ, rep


def run_ragged(args) -> dict:
    """Ragged workload: 262,144 trajectories, 60% LEO (alt U(300,2000) km),
    25% MEO (a U(20,000, 30,000) km), 15% GEO (a = 42,164 km, e < 1e-3),
    durations log-uniform 1 h .. 30 d, degree/order 20, RK8(7)13M rtol 1e-11,
    gravity only.  One persistent launch under a HIP graph; launch shape from
    the autotuner's pilot."""
    import bqp_amd
    from bqp_amd import inputs
    from bqp_amd.autotune import autotune
    from bqp_amd.device import DeviceArray, synchronize

    n = args.rows if args.rows != 65536 else 262144
    w = inputs.workload("w2", n=n)
    field = bqp_amd.kaula(20)
    y0 = DeviceArray.from_numpy(w["y0"])
    t0 = DeviceArray.from_numpy(np.zeros(n))
    tf = DeviceArray.from_numpy(w["tf"])
    orbit_days = float(np.sum(w["tf"]) / 86400.0)
    prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-11)
    tune = autotune(prop, w["y0"], 0.0, w["tf"], use_cache=False, cache_path=None)
    prop.set_launch(**{**tune["launch"], "use_graph": 1})
    frhs = prop.flops_per_rhs
    fstage = stage_flops("rk87")
    yf = DeviceArray((n, 6))
    status = DeviceArray((n,), np.int32)
    for _ in range(max(args.warmup, 1)):  # the first call captures the graph
        prop.propagate_batch_final_dev(y0, t0, tf, None, yf, status)
    wall = kern = flops = 0.0
    for _ in range(args.steps):
        synchronize()
        ts = time.perf_counter()
        _, _, st = prop.propagate_batch_final_dev(y0, t0, tf, None, yf, status)
        synchronize()
        wall += time.perf_counter() - ts
        kern += st["kernel_ms"]
        flops += st["evals"] * frhs + (st["steps"] + st["rejects"]) * fstage
    ms = wall / args.steps * 1e3
    kms = kern / args.steps
    ok = int(np.sum(status.numpy() == 0))
    prop.close()
    achieved = flops / args.steps / (kms * 1e-3) / 1e12
    print(f"[bench] ragged: {orbit_days / (ms * 1e-3):.1f} orbit-days/s, {ms:.1f} ms/step, {achieved:.2f} TFLOP/s, "
          f"ok {ok}/{n}, launch {tune['launch']}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline_ragged(w, field, threads)
    return {
        "metric": "orbit-days/s, 262,144 mixed LEO/MEO/GEO, log-uniform 1 h-30 d, deg-20 (ragged batch)",
        "value": orbit_days / (ms * 1e-3),
        "unit": "orbit-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded mixed-orbit elements and durations (PCG64 seed 0), seeded Kaula degree-20 field",
        "config": {"workload": f"{n} mixed orbits, durations log-uniform 1 h..30 d ({orbit_days:.0f} orbit-days), "
                               "RK8(7)13M rtol 1e-11, degree/order 20, gravity only, one persistent launch under a "
                               "HIP graph", "orbits": n, "degree": 20, "launch": tune["launch"],
                   "autotune_trials": tune["trials"]},
        "roofline": {"bound": "mfma", "pipe": "FP64 vector ALU", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel_ms": kms,
                     "flops_per_rhs": frhs},
        "cpu_baseline": cpu,
    }


def run_surrogate_study(args) -> dict:
    """The learned-surrogate form on the headline workload (65,536 LEO x 7 d,
    degree 70, Vallado drag, RK8(7)13M rtol 1e-12): device training on 10M
    residual samples, MFMA inference rates (BF16 and FP8) against their
    rooflines, network error against the 1e-9 m/s^2 check bound, the checked
    in-integrator form against plain FP64 on the full workload, and the
    unchecked network (an ablation: no exact fallback) on a 1-hour arc."""
    import bqp_amd
    from bqp_amd import inputs
    from bqp_amd.device import DeviceArray, synchronize

    n = args.rows
    days = args.days
    w = inputs.workload("w1", n=n)
    field = bqp_amd.kaula(70)
    prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado")
    sur, feat, targ, train_rep = train_surrogate(prop, args.sur_rows, args.sur_steps)
    scale = sur.out_scale
    vfeat, vtarg, _ = sur.dataset(1_000_000, seed=977)
    T = vtarg.numpy()
    res = np.linalg.norm(T * scale, axis=1)
    infer = {}
    for fmt in ("fp32", "bf16", "fp8"):
        if fmt != "fp32":
            sur.pack(fmt, calib=feat.rows(0, 65536))
        o, _ = sur.infer(vfeat, fmt)
        err = np.linalg.norm((o.numpy() - T) * scale, axis=1)
        rec = {"err_rms_m_s2": float(np.sqrt(np.mean(err ** 2))), "err_median_m_s2": float(np.median(err)),
               "err_max_m_s2": float(err.max()), "frac_within_1e-9": float(np.mean(err <= 1e-9))}
        if fmt != "fp32":
            _, ms = sur.infer(feat, fmt)
            _, ms = sur.infer(feat, fmt)
            tf_s = MLP_FLOPS * feat.shape[0] / (ms * 1e-3) / 1e12
            peak = BF16_PEAK_TFLOPS if fmt == "bf16" else FP8_PEAK_TFLOPS
            rec.update({"rows": int(feat.shape[0]), "kernel_ms": ms, "achieved_tflops": tf_s, "peak_tflops": peak,
                        "frac": tf_s / peak})
        infer[fmt] = rec
        print(f"[bench] surrogate {fmt}: {rec}", file=sys.stderr, flush=True)
    residual = {"rms_m_s2": float(np.sqrt(np.mean(res ** 2))), "max_m_s2": float(res.max())}

    y0 = DeviceArray.from_numpy(w["y0"])
    t0 = DeviceArray.from_numpy(np.zeros(n))
    bc = DeviceArray.from_numpy(w["bc"])
    fstage = stage_flops("rk87")

    def run(span, warm):
        tf = DeviceArray.from_numpy(np.full(n, span))
        yf = DeviceArray((n, 6))
        st_ = DeviceArray((n,), np.int32)
        if warm:
            prop.propagate_batch_final_dev(y0, t0, tf, bc, yf, st_)
        synchronize()
        ts = time.perf_counter()
        _, _, st = prop.propagate_batch_final_dev(y0, t0, tf, bc, yf, st_)
        synchronize()
        wall = time.perf_counter() - ts
        return yf.numpy(), st_.numpy(), wall, st

    out_runs = {}
    spans = {"full": days * 86400.0, "1h": 3600.0}
    ref = {}
    for key, span in spans.items():
        prop.reconfigure(precision="fp64")
        y, s, wall, st = run(span, warm=(key == "full"))
        ref[key] = y
        out_runs[f"fp64_{key}"] = {"orbit_days_s": n * span / 86400.0 / wall, "wall_s": wall, "evals": st["evals"],
                                   "steps": st["steps"], "ok": int(np.sum(s == 0))}
        print(f"[bench] fp64 {key}: {out_runs[f'fp64_{key}']}", file=sys.stderr, flush=True)
    prop.reconfigure(precision="surrogate")
    cases = [("bf16", args.sur_check, "full"), ("fp8", args.sur_check, "full"), ("bf16", 0, "1h"), ("fp8", 0, "1h")]
    for fmt, ce, key in cases:
        sur.pack(fmt, calib=feat.rows(0, 65536))
        sur.set_check(ce, 1e-9)
        y, s, wall, st = run(spans[key], warm=False)
        d = np.nan_to_num(np.linalg.norm(y[:, :3] - ref[key][:, :3], axis=1), nan=np.inf)
        base = out_runs[f"fp64_{key}"]
        rec = {"orbit_days_s": n * spans[key] / 86400.0 / wall, "wall_s": wall,
               "speedup_vs_fp64": base["wall_s"] / wall, "evals": st["evals"], "steps": st["steps"],
               "evals_vs_fp64": st["evals"] / max(base["evals"], 1), "ok": int(np.sum(s == 0)),
               "pos_err_median_m": float(np.median(d)), "pos_err_max_m": float(d.max()),
               "stats": sur.last_stats(),
               "mlp_tflops": sur.last_stats()["net_evals"] * MLP_FLOPS / (st["kernel_ms"] * 1e-3) / 1e12}
        tag = f"surrogate_{fmt}_{'checked' if ce else 'unchecked'}_{key}"
        out_runs[tag] = rec
        print(f"[bench] {tag}: {rec}", file=sys.stderr, flush=True)
    prop.close()
    main_run = out_runs["surrogate_bf16_checked_full"]
    achieved = infer["bf16"]["achieved_tflops"]
    return {
        "metric": METRIC,
        "value": main_run["orbit_days_s"],
        "unit": "orbit-days/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": main_run["wall_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network / f64 state",
        "data": "synthetic: seeded LEO elements (PCG64 seed 0), seeded Kaula degree-70 field, seeded residual samples",
        "config": {"workload": f"{n} LEO orbits x {days:g} days, RK8(7)13M rtol 1e-12, degree/order 70, Vallado "
                               f"drag; learned residual above degree 4 with exact fallback (check every "
                               f"{args.sur_check} evaluations, bound 1e-9 m/s^2)",
                   "form": "surrogate (checked)", "truth_degree": 70},
        "roofline": {"bound": "mfma", "pipe": "BF16 MFMA (standalone network kernel, 10M rows)",
                     "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": None,
        "training": train_rep,
        "residual": residual,
        "inference": infer,
        "runs": out_runs,
    }


def _round_bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even) -> float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def run_floors(args) -> dict:
    """Error floors of the two reduced-precision forms on the headline field
    (Kaula degree 70, LEO band 6,578-8,378 km, RK8(7)13M rtol 1e-12).

    Learned surrogate: the network trained from scratch for each step budget
    in --floor-steps (10M samples, batch 8,192, cosine learning rate); on 1M
    held-out samples, in m/s^2: the FP32 network against the exact residual
    (the fitting floor), the FP32 network on BF16-rounded features against
    the same network on FP32 features (the input-quantisation floor, which
    the BF16 and FP8 forms share: both read BF16 features), and the BF16 /
    FP8 MFMA networks against the FP32 network (all quantisation together).
    Then the network trained on degree-120 truth (the second budget): the
    size of degrees 71-120 at the same positions, and its error against
    degree 120 and against the degree-70 field the propagation uses.

    Split form: on one accepted-size step of each of 4,096 rows (the mean
    step of that row's 1-day FP64 run), the change of the embedded error
    estimate when every stage acceleration comes from the split-precision
    gravity instead of the FP64 one, in units of the WRMS error norm
    (scale atol + rtol |y| per component, as the kernel's controller), next
    to the FP64 estimate itself; and the RHS evaluations of the split form
    over the same day against FP64's."""
    import bqp_amd
    from bqp_amd import inputs
    from bqp_amd.device import DeviceArray, synchronize
    from bqp_amd.propagator import tableau

    field = bqp_amd.kaula(70)
    prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-12)
    out = {"metric": "surrogate and split-form error floors", "field": "Kaula degree/order 70 (seed 0)"}

    # ------------------------------------------------------------ surrogate
    budgets = [int(s) for s in args.floor_steps.split(",") if s]
    sur = bqp_amd.Surrogate(prop, nlow=4, r_lo=6578.0e3, r_hi=8378.0e3, seed=0)
    feat, targ, _ = sur.dataset(args.sur_rows, seed=1)
    vfeat, vtarg, _ = sur.dataset(1_000_000, seed=977)
    T = vtarg.numpy()
    vfeat_b = DeviceArray.from_numpy(_round_bf16(vfeat.numpy()))
    calib = feat.rows(0, 65536)
    curve = []
    for steps in budgets:
        sur = bqp_amd.Surrogate(prop, nlow=4, r_lo=6578.0e3, r_hi=8378.0e3, seed=0)
        scale = sur.out_scale
        t = time.perf_counter()
        losses = sur.fit(feat, targ, steps=steps, batch=8192, lr_max=2e-3, lr_min=1e-5, chunk=500, seed=0)
        synchronize()
        train_s = time.perf_counter() - t
        o32 = sur.infer(vfeat, "fp32")[0].numpy()
        o32b = sur.infer(vfeat_b, "fp32")[0].numpy()
        sur.pack("bf16")
        obf = sur.infer(vfeat, "bf16")[0].numpy()
        sur.pack("fp8", calib=calib)
        o8 = sur.infer(vfeat, "fp8")[0].numpy()

        def rms(d):
            e = np.linalg.norm(d * scale, axis=1)
            return {"rms": float(np.sqrt(np.mean(e ** 2))), "median": float(np.median(e)), "max": float(e.max()),
                    "frac_le_1e-9": float(np.mean(e <= 1e-9))}

        rec = {"steps": steps, "train_s": train_s, "loss_last": float(np.mean(losses[-50:])),
               "fit_fp32_vs_truth": rms(o32 - T), "input_bf16_vs_fp32": rms(o32b - o32),
               "bf16_vs_fp32": rms(obf - o32), "fp8_vs_fp32": rms(o8 - o32), "bf16_vs_truth": rms(obf - T)}
        curve.append(rec)
        print(f"[bench] floors: {steps} steps ({train_s:.1f} s): fit {rec['fit_fp32_vs_truth']['rms']:.3e}, "
              f"input bf16 {rec['input_bf16_vs_fp32']['rms']:.3e}, bf16 {rec['bf16_vs_fp32']['rms']:.3e}, "
              f"fp8 {rec['fp8_vs_fp32']['rms']:.3e} m/s^2 RMS", file=sys.stderr, flush=True)
    scale70 = sur.out_scale
    res = np.linalg.norm(T * scale70, axis=1)
    out["surrogate"] = {"residual_rms_m_s2": float(np.sqrt(np.mean(res ** 2))), "check_bound_m_s2": 1e-9,
                        "relative_accuracy_needed": 1e-9 / float(np.sqrt(np.mean(res ** 2))),
                        "budgets": curve}

    # Training truth of degree 120 (the Kaula field of degree 70 is a prefix of
    # the degree-120 one): the network learns degrees 5..120 while the
    # propagation field, its exact fallback and its check stop at 70.
    steps120 = budgets[min(1, len(budgets) - 1)]
    p120 = bqp_amd.Propagator(bqp_amd.kaula(120), method="rk87", rtol=1e-12)
    s120 = bqp_amd.Surrogate(p120, nlow=4, r_lo=6578.0e3, r_hi=8378.0e3, seed=0)
    f120, t120, _ = s120.dataset(args.sur_rows, seed=1)
    v120f, v120t, _ = s120.dataset(1_000_000, seed=977)
    same_positions = bool(np.array_equal(v120f.numpy(), vfeat.numpy()))
    s120.fit(f120, t120, steps=steps120, batch=8192, lr_max=2e-3, lr_min=1e-5, chunk=500, seed=0)
    o120 = s120.infer(v120f, "fp32")[0].numpy() * s120.out_scale
    T120 = v120t.numpy() * s120.out_scale
    T70 = T * scale70

    def stat(d):
        e = np.linalg.norm(d, axis=1)
        return {"rms": float(np.sqrt(np.mean(e ** 2))), "median": float(np.median(e)), "max": float(e.max())}

    out["surrogate"]["truth_degree_120"] = {
        "steps": steps120, "same_positions": same_positions,
        "degrees_71_120_m_s2": stat(T120 - T70),
        "fit_fp32_vs_degree_120": stat(o120 - T120),
        "fp32_vs_degree_70_field": stat(o120 - T70)}
    print(f"[bench] floors: degree-120 truth ({steps120} steps): degrees 71-120 part "
          f"{stat(T120 - T70)['rms']:.3e}, fit vs 120 {stat(o120 - T120)['rms']:.3e}, vs the degree-70 field "
          f"{stat(o120 - T70)['rms']:.3e} m/s^2 RMS", file=sys.stderr, flush=True)
    p120.close()

    # ------------------------------------------------------------ split form
    n = 4096
    w = inputs.workload("w1", n=n)
    tb = tableau("rk87")
    S, A, c, e = tb["stages"], tb["a"], tb["c"], tb["e"]
    evals = {}
    hmean = None
    for prec in ("fp64", "split"):
        with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, precision=prec) as p:
            _, st, info = p.propagate_batch_final(w["y0"], 0.0, np.full(n, 86400.0))
            evals[prec] = int(np.sum(info["evals_per_row"]))
            if prec == "fp64":
                hmean = 86400.0 / np.maximum(info["steps_per_row"], 1)
    cfg = prop.config
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-12, precision="split") as ps:
        y = w["y0"].copy()
        h = hmean[:, None]
        K64 = np.zeros((S, n, 6))
        dA = np.zeros((S, n, 3))
        for s in range(S):
            ys = y + h * np.einsum("j,jnk->nk", A[s, :s], K64[:s]) if s else y.copy()
            a64 = prop.gravity_accel(ys[:, :3])
            a32 = ps.gravity_accel(ys[:, :3])
            K64[s, :, :3] = ys[:, 3:]
            K64[s, :, 3:] = a64
            dA[s] = a32 - a64
    err64 = h * np.einsum("j,jnk->nk", e, K64)
    # the stage velocities inherit the acceleration noise: V_s = v + h sum_j a_sj A_j
    dK = np.zeros((S, n, 6))
    for s in range(S):
        dv = h * np.einsum("j,jnk->nk", A[s, :s], dA[:s]) if s else np.zeros((n, 3))
        dK[s, :, :3] = dv
        dK[s, :, 3:] = dA[s]
    derr = h * np.einsum("j,jnk->nk", e, dK)
    atol = np.array([cfg.atol_pos] * 3 + [cfg.atol_vel] * 3)
    sc = atol + cfg.rtol * np.abs(y)  # the step's end state differs by O(h |y'|) -- immaterial here
    wrms = lambda d: np.sqrt(np.mean((d / sc) ** 2, axis=1))
    en64, enoise = wrms(err64), wrms(derr)
    dA_norm = np.linalg.norm(dA, axis=2)
    out["split"] = {"rows": n, "rtol": cfg.rtol, "atol_pos_m": cfg.atol_pos, "atol_vel_m_s": cfg.atol_vel,
                    "mean_step_s": {"median": float(np.median(hmean)), "min": float(hmean.min()),
                                    "max": float(hmean.max())},
                    "gravity_noise_m_s2": {"rms": float(np.sqrt(np.mean(dA_norm ** 2))), "max": float(dA_norm.max())},
                    "en_fp64": {"median": float(np.median(en64)), "p90": float(np.percentile(en64, 90))},
                    "en_noise": {"median": float(np.median(enoise)), "p90": float(np.percentile(enoise, 90)),
                                 "max": float(enoise.max())},
                    "vel_noise_over_scale_median": float(np.median(np.abs(derr[:, 3:]) / sc[:, 3:])),
                    "pos_noise_over_scale_median": float(np.median(np.abs(derr[:, :3]) / sc[:, :3])),
                    "evals_1day": evals, "evals_ratio": evals["split"] / evals["fp64"]}
    print(f"[bench] floors: split noise en median {np.median(enoise):.3g} (fp64 en median {np.median(en64):.3g}), "
          f"evals ratio {evals['split'] / evals['fp64']:.3f}", file=sys.stderr, flush=True)
    prop.close()
    return out


def run_population(args) -> dict:
    """Population evaluation: 32 NSGA-II generations x 4,096 candidates
    (131,072 arcs) over seeded LEO departure states with dv U(-50, 50) m/s
    per axis and coast U(0.5, 10) d, degree/order 50 + Vallado drag,
    RK8(7)13M rtol 1e-12, ground event 100 km above the WGS84 ellipsoid (a
    candidate that re-enters ends there and loses; its arc is not counted in
    the orbit-days).  Each generation is one ragged device batch; the field
    and the surrogate stay resident between generations.  Forms: plain FP64,
    then the checked surrogate with a BF16 and with an FP8 network."""
    import bqp_amd
    from bqp_amd import inputs
    from bqp_amd.population import DepartureProblem, Nsga2

    pop, gens = args.pop, args.generations
    w = inputs.workload("w1", n=pop, seed=7)
    field = bqp_amd.kaula(50)
    r_target = 6378137.0 + 800e3
    forms = [f for f in args.forms.split(",") if f] if args.forms != "fp64,split,surrogate,encke" else \
        ["fp64", "surrogate_bf16", "surrogate_fp8"]
    out, finals, sur = {}, {}, None
    # the reference's ground event: a candidate that re-enters fails there
    # (two_phase_transfer_rs/src/evaluate.rs:1044-1060) and loses
    prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado",
                              r_min=bqp_amd.R_EARTH + bqp_amd.GROUND_ALTITUDE_M,
                              ground_flattening=bqp_amd.WGS84_FLATTENING)
    if args.launch:
        prop.set_launch(**json.loads(args.launch))
    for form in forms:
        if form.startswith("surrogate"):
            if sur is None:
                prop.reconfigure(precision="fp64")
                sur, feat, _, train_rep = train_surrogate(prop, args.sur_rows, args.sur_steps)
            fmt = form.split("_")[1]
            sur.pack(fmt, calib=feat.rows(0, 65536))
            sur.set_check(args.sur_check, 1e-9)
            prop.reconfigure(precision="surrogate")
        else:
            prop.reconfigure(precision=form)
        prob = DepartureProblem(prop, w["y0"], w["bc"], r_target)
        rng = np.random.Generator(np.random.PCG64(11))
        t = time.perf_counter()
        opt = Nsga2(prob, prob.lower, prob.upper, pop_size=pop, seed=11,
                    initial=prob.initial_population(pop, rng))
        first = prob.last[0].copy()

        def progress(o):
            print(f"[bench] population {form}: generation {o.generation + 1}/{gens}, device "
                  f"{prob.device_s:.1f} s, {prob.orbit_days:.0f} orbit-days", file=sys.stderr, flush=True)

        opt.run(gens - 1, callback=progress)
        wall = time.perf_counter() - t
        Xf, Ff = opt.front()
        rec = {"orbit_days_s": prob.orbit_days / prob.device_s, "orbit_days": prob.orbit_days,
               "device_s": prob.device_s, "wall_s": wall, "host_share": 1.0 - prob.device_s / wall,
               "evaluations": opt.evaluations, "generations": opt.generation + 1, "rhs_evals": prob.evals,
               "front_size": int(Xf.shape[0]), "front_min_dv_m_s": float(Ff[:, 0].min()),
               "front_min_miss_km": float(Ff[:, 1].min())}
        if form.startswith("surrogate"):
            rec["surrogate_stats_last_generation"] = sur.last_stats()
        finals[form] = first
        out[form] = rec
        print(f"[bench] population {form}: {rec}", file=sys.stderr, flush=True)
    prop.close()
    for form in forms:
        if form != "fp64" and "fp64" in finals:
            d = np.nan_to_num(np.linalg.norm(finals[form][:, :3] - finals["fp64"][:, :3], axis=1), nan=np.inf)
            out[form].update({"gen0_pos_err_median_m": float(np.median(d)), "gen0_pos_err_max_m": float(d.max()),
                              "speedup_vs_fp64": out["fp64"]["device_s"] / out[form]["device_s"]})
    base = out.get("fp64") or next(iter(out.values()))
    return {
        "metric": "orbit-days/s, NSGA-II population evaluation: 32 generations x 4,096 arcs, deg-50 + drag",
        "value": base["orbit_days_s"],
        "unit": "orbit-days/s",
        "n_gpus": 1,
        "steps": gens,
        "warmup": 0,
        "ms_per_step": base["device_s"] / gens * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded LEO departures (PCG64 seed 7), dv U(-50,50) m/s per axis, coast U(0.5,10) d, "
                "seeded Kaula degree-50 field",
        "config": {"workload": f"{gens} NSGA-II generations x {pop} candidates, objectives |dv| and final-radius "
                               f"miss vs {r_target / 1e3:.0f} km, RK8(7)13M rtol 1e-12, degree/order 50, Vallado "
                               "drag; one ragged device batch per generation", "forms": forms},
        "roofline": None,
        "cpu_baseline": None,
        "forms": out,
    }


def cpu_baseline_ragged(w, field, threads: int) -> dict:
    from oracle import oracle as orc

    f = orc.Field(field.cbar, field.sbar, field.nmax, field.mu, field.radius)
    model = orc.make_model(f)
    integ = orc.make_integ(method="rk87", rtol=1e-11, canonical=(field.radius, field.mu))
    rate_total, parts = 0.0, []
    n = w["y0"].shape[0]
    for cls, frac in ((0, 0.60), (1, 0.25), (2, 0.15)):
        members = np.nonzero(w["orbit_class"] == cls)[0]
        rows = 4 * threads
        while True:  # grow the sample until it is a few seconds of CPU work
            idx = members[:rows]
            t0 = time.perf_counter()
            _, st, _, _ = orc.propagate_batch(model, integ, w["y0"][idx], 0.0, w["tf"][idx], threads=threads)
            wall = time.perf_counter() - t0
            if wall >= 3.0 or rows >= members.size:
                break
            rows = min(members.size, int(rows * max(2.0, 3.5 / max(wall, 1e-3))))
        days = float(np.sum(w["tf"][idx]) / 86400.0)
        # time per orbit-day of this class, weighted by the class's share of the workload's orbit-days
        share = float(np.sum(w["tf"][w["orbit_class"] == cls]) / np.sum(w["tf"]))
        rate_total += share * wall / days
        parts.append(f"class {cls}: {idx.size} rows, {days:.1f} orbit-days in {wall:.1f} s")
    return {"value": 1.0 / rate_total, "unit": "orbit-days/s", "cores": threads, "kind": "port",
            "sample": "first rows of each orbit class at full duration (at least 4 x threads rows, grown until the "
                      "class takes >= 3 s), weighted by each class's share of the workload's orbit-days; "
                      + "; ".join(parts)}


def cpu_threads() -> int:
    try:
        return min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        return min(16, os.cpu_count() or 1)


def run_form(form, args, field, bufs, n, warm_passes):
    """One form of the headline workload: `warm_passes` untimed passes (a
    one-hour pass when warm_passes is 0 and --warmup is set), then
    args.steps or 1 timed passes over the whole batch, each bracketed by a
    device synchronisation.  Returns (result record, final states)."""
    import bqp_amd
    from bqp_amd.device import DeviceArray, synchronize

    y0, t0, tf, tf_short, bc = bufs
    if form == "encke":  # the reference's formulation: deviation from a rebased two-body baseline
        prop = bqp_amd.Propagator(field, method="rk87", control="absolute", eps=args.encke_eps,
                                  drag="vallado", formulation="encke")
    else:
        prop = bqp_amd.Propagator(field, method="rk87", rtol=1e-12, drag="vallado",
                                  precision="surrogate" if form == "surrogate" else form)
    if args.launch:
        prop.set_launch(**json.loads(args.launch))
    sur = sur_report = None
    if form == "surrogate":
        sur, _, _, sur_report = train_surrogate(prop, args.sur_rows, args.sur_steps)
        sur.pack("bf16")
        sur.set_check(args.sur_check, 1e-9)
    frhs = prop.flops_per_rhs
    fstage = stage_flops("rk87")
    yf = DeviceArray((n, 6))
    status = DeviceArray((n,), np.int32)
    steps_t = DeviceArray((n,), np.int32)
    evals_t = DeviceArray((n,), np.int32)
    for i in range(warm_passes):
        prop.propagate_batch_final_dev(y0, t0, tf, bc, yf, status, steps_t, evals_t)
        print(f"[bench] {form}: warm-up pass {i + 1}/{warm_passes} done", file=sys.stderr, flush=True)
    if warm_passes == 0 and args.warmup:
        # a one-hour pass brings clocks and caches up without a full extra pass
        prop.propagate_batch_final_dev(y0, t0, tf_short, bc, yf, status, steps_t, evals_t)
    timed = args.steps if warm_passes else 1
    synchronize()
    wall = kern_ms = flops = 0.0
    for i in range(timed):
        synchronize()
        ts = time.perf_counter()
        _, _, st = prop.propagate_batch_final_dev(y0, t0, tf, bc, yf, status, steps_t, evals_t)
        synchronize()
        wall += time.perf_counter() - ts
        kern_ms += st["kernel_ms"]
        flops += (st["flops"] if form == "surrogate" else st["evals"] * frhs) + \
            (st["steps"] + st["rejects"]) * fstage
        print(f"[bench] {form}: timed pass {i + 1}/{timed}: {st['kernel_ms']:.1f} ms kernel", file=sys.stderr,
              flush=True)
    ok = int(np.sum(status.numpy() == 0))
    ev_rows = evals_t.numpy()
    ms = wall / timed * 1e3
    kms = kern_ms / timed
    achieved = flops / timed / (kms * 1e-3) / 1e12
    rec = {
        "value": n * args.days / (ms * 1e-3),
        "timed_passes": timed,
        "ms_per_step": ms,
        "kernel_ms": kms,
        "rows_ok": ok,
        "rhs_evals": int(st["evals"]),
        "steps": int(st["steps"]),
        "rejects": int(st["rejects"]),
        "flops_per_rhs": frhs,
        "achieved_tflops": achieved,
        "launch": getattr(prop, "launch", None),
        # the persistent kernel and shape the timed passes ran with
        "launch_used": prop.last_launch(),
        # per-row work spread: every lane holds one row, so the rows with
        # the most evaluations set the launch's length
        "evals_per_row": {"min": int(ev_rows.min()), "mean": float(ev_rows.mean()), "max": int(ev_rows.max())},
    }
    if form == "surrogate":
        rec["surrogate_stats"] = sur.last_stats()
        rec["training"] = sur_report
    final = yf.numpy()
    prop.close()
    print(f"[bench] {form}: {rec['value']:.1f} orbit-days/s, {ms:.1f} ms/step, kernel {kms:.1f} ms, "
          f"{achieved:.2f} TFLOP/s, ok {ok}/{n}", file=sys.stderr, flush=True)
    return rec, final


# Cost of one timed pass of each later form, as a multiple of the FP64
# pass, plus its fixed set-up seconds (surrogate: device training), used to
# keep the whole run inside --budget-s.
FORM_COST = {"split": (2.3, 2.0), "surrogate": (1.15, 20.0), "encke": (1.3, 2.0), "fp64": (1.1, 2.0)}


def main():
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=65536)
    ap.add_argument("--days", type=float, default=7.0)
    ap.add_argument("--forms", default="fp64,split,surrogate,encke")
    ap.add_argument("--budget-s", type=float, default=450.0,
                    help="wall-clock budget of the whole run: forms after the headline are skipped when "
                         "their projected time would cross it")
    ap.add_argument("--encke-eps", type=float, default=1e-6,
                    help="encke form: absolute error bound on the deviation (m, m/s) per step")
    ap.add_argument("--sur-rows", type=int, default=10_000_000, help="surrogate training samples")
    ap.add_argument("--sur-steps", type=int, default=3000, help="surrogate Adam steps (batch 8192)")
    ap.add_argument("--sur-check", type=int, default=100, help="surrogate: check every k-th evaluation")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU time of the baseline sample")
    ap.add_argument("--launch", default="", help='JSON launch shape, e.g. {"waves_per_group": 2}')
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (a known-size kernel for "
                         "calibrating HBM byte counters)")
    ap.add_argument("--pop", type=int, default=4096, help="w4: population size")
    ap.add_argument("--generations", type=int, default=32, help="w4: generations (incl. the initial one)")
    ap.add_argument("--floor-steps", default="1000,3000,10000,30000",
                    help="floors: training step budgets of the surrogate's error-floor curve")
    ap.add_argument("--workload", default="w1", choices=["w1", "w2", "w3", "w4", "floors"],
                    help="w1: 65,536 LEO x 7 d deg-70 (the headline); w2: ragged 262,144 mixed orbits deg-20; "
                         "w3: the surrogate study (training, MFMA inference rooflines, fallback, ablation); "
                         "w4: NSGA-II population evaluation (32 x 4,096 arcs, deg-50 + drag); "
                         "floors: error floors of the surrogate (fitting, input and weight quantisation) and of "
                         "the split form's error estimate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the propagation path runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; trajectories would shard by cost-weighted "
              f"index ranges, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    if args.workload == "w2":
        print(json.dumps(run_ragged(args)))
        return
    if args.workload == "w3":
        print(json.dumps(run_surrogate_study(args)))
        return
    if args.workload == "w4":
        print(json.dumps(run_population(args)))
        return
    if args.workload == "floors":
        print(json.dumps(run_floors(args)))
        return

    import bqp_amd
    from bqp_amd import inputs
    from bqp_amd.device import DeviceArray

    n = args.rows
    w = inputs.workload("w1", n=n)
    field = bqp_amd.kaula(70)
    tf_val = args.days * 86400.0
    forms = [f for f in args.forms.split(",") if f]
    if not forms:
        print("--forms names no form", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs first, before any GPU work is timed
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, field, cpu_threads(), args.cpu_seconds)
        print(f"[bench] cpu baseline {cpu['value']:.3f} orbit-days/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    # inputs resident in HBM before any timed region
    bufs = (DeviceArray.from_numpy(w["y0"]), DeviceArray.from_numpy(np.zeros(n)),
            DeviceArray.from_numpy(np.full(n, tf_val)), DeviceArray.from_numpy(np.full(n, min(tf_val, 3600.0))),
            DeviceArray.from_numpy(w["bc"]))

    # the headline form: --warmup full passes, then --steps timed passes
    head = forms[0]
    base, head_final = run_form(head, args, field, bufs, n, args.warmup)
    line = {
        "metric": METRIC,
        "value": base["value"],
        "unit": "orbit-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": base["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head in ("fp64", "encke") else ("f64+f32" if head == "split" else "f64+bf16"),
        "data": "synthetic: seeded LEO elements (PCG64 seed 0) and seeded Kaula degree-70 field",
        "config": {
            "workload": f"{n} LEO orbits x {args.days:g} days, RK8(7)13M rtol 1e-12, degree/order 70, "
                        f"Vallado exponential drag",
            "orbits": n,
            "days": args.days,
            "degree": 70,
            "method": "RK8(7)13M",
            "rtol": 1e-12,
            "drag": "vallado",
            "form": head,
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "FP64 vector ALU (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s)",
            "achieved": base["achieved_tflops"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": base["achieved_tflops"] / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": kernel_name(base.get("launch_used"), head),
            "kernel_ms": base["kernel_ms"],
            "flops_per_rhs": base["flops_per_rhs"],
            "rhs_evals": base["rhs_evals"],
            "evals_per_row": base["evals_per_row"],
        },
        "cpu_baseline": cpu,
    }
    traffic = committed_traffic() if (n == 65536 and args.days == 7.0 and head == "fp64") else None
    if traffic is not None:
        # HBM bytes per launch cannot be counted inside this process (the
        # counters need rocprofv3 --pmc); the figure is the committed one
        line["roofline"]["traffic"] = traffic["bytes"]
        line["roofline"]["traffic_source"] = traffic["source"]
    print(json.dumps(line), flush=True)

    # the other forms: one timed pass each, inside the wall-clock budget
    forms_out, skipped = {}, {}
    for form in forms[1:]:
        factor, setup = FORM_COST.get(form, (2.5, 5.0))
        projected = setup + (factor * base["ms_per_step"] + 3.6e6 / (args.days * 86400.0) * base["ms_per_step"]) * 1e-3
        elapsed = time.perf_counter() - t_start
        if elapsed + projected > args.budget_s:
            skipped[form] = f"skipped: {elapsed:.0f} s elapsed + {projected:.0f} s projected > {args.budget_s:.0f} s"
            print(f"[bench] {form}: {skipped[form]}", file=sys.stderr, flush=True)
            continue
        rec, final = run_form(form, args, field, bufs, n, 0)
        d = np.nan_to_num(np.linalg.norm(final[:, :3] - head_final[:, :3], axis=1), nan=np.inf)
        out = {
            "value": rec["value"],
            f"speedup_vs_{head}": rec["value"] / base["value"],
            f"max_pos_err_m_vs_{head}": float(d.max()),
            f"median_pos_err_m_vs_{head}": float(np.median(d)),
            "achieved_tflops": rec["achieved_tflops"],
            "kernel_ms": rec["kernel_ms"],
            "timed_passes": rec["timed_passes"],
            "rows_ok": rec["rows_ok"],
            "rhs_evals": rec["rhs_evals"],
        }
        if form == "encke":
            out.update({"formulation": "Encke deviation on a two-body baseline, Battin correction, rebased every "
                                       "5,400 s (fresh controller per segment), absolute error control",
                        "eps": args.encke_eps, "steps": rec["steps"]})
        if form == "surrogate":
            out.update({"network": "16 features -> 4 x 256 SiLU -> 3, BF16 MFMA, FP32 accumulate, residual above "
                                   "deg 4", "check_every": args.sur_check, "threshold_m_s2": 1e-9,
                        "stats": rec["surrogate_stats"], "training": rec["training"]})
        forms_out[form] = out
    if len(forms) > 1:
        line["forms"] = forms_out
        if skipped:
            line["forms_skipped"] = skipped
        line["wall_s"] = time.perf_counter() - t_start
        print(json.dumps(line), flush=True)
    if args.copy_mib:
        copy_calibration(args.copy_mib)


def copy_calibration(mib: int):
    """One device-to-device copy of `mib` MiB: reads and writes exactly that
    many bytes, the reference for scaling HBM byte counters."""
    import ctypes as C

    from bqp_amd.device import D2D, DeviceArray, hip, synchronize

    n = (mib << 20) // 8
    a = DeviceArray((n,))
    b = DeviceArray((n,))
    synchronize()
    rc = hip().hipMemcpy(C.c_void_p(b.ptr), C.c_void_p(a.ptr), a.nbytes, D2D)
    synchronize()
    print(f"[bench] calibration copy of {mib} MiB ({a.nbytes} bytes), rc {rc}", file=sys.stderr, flush=True)


def kernel_name(used, form: str) -> str:
    """Name of the persistent kernel a call ran, as the kernel trace shows it."""
    prec = {"fp64": "0", "split": "1", "encke": "4"}.get(form, "0")
    if not used:
        return "bqp::propagate_kernel<2, 0, G>"
    if used["kernel"] == "pair":
        return f"bqp::propagate_pair_kernel<2, {prec}> (two 64-trajectory groups per 8-wave workgroup)"
    if used["kernel"] == "team":
        return f"bqp::propagate_team_kernel<2, {prec}> ({used['teams']} groups per 16-wave workgroup)"
    return f"bqp::propagate_kernel<2, {prec}, {used['waves_per_group']}>"


def committed_traffic():
    """HBM bytes per launch of the propagation kernel on the headline
    workload, from the committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel's calibration); None if absent."""
    rel = os.path.join("profiles", "propagate_rk87_deg70_leo65536_7d_traffic.json")
    try:
        with open(os.path.join(ROOT, rel)) as fh:
            return {"bytes": float(json.load(fh)["hbm_bytes_per_launch"]),
                    "source": f"{rel} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, not this run)"}
    except (OSError, ValueError, KeyError, TypeError):
        return None


if __name__ == "__main__":
    main()
