#!/usr/bin/env python3
"""Throughput of the adaptive-rank TLR log-likelihood plus gradient on one MI355X, in reduced-precision
arithmetic beside plain FP64.

Workload: N = 40,000 sites i.i.d. uniform in the unit square (default_rng(0)), Morton-ordered, Matern
nu=3/2, tiles b = 500 (80 x 80 tile grid), adaptive rank rule tau = 0.9, ranks re-chosen at each of 20 theta
on a Latin hypercube (scipy.stats.qmc.LatinHypercube(d=3, seed=0)) over sigma2 in [0.5, 2], range in
[0.1, 0.6], eta in [0.01, 0.1].  One step = one batched evaluation of log L and its gradient in log theta at
those 20 theta (log L, log-det, quadratic form, per-tile ranks; gradient at frozen ranks and frozen
arithmetic: the sigma2 component in closed form, range and eta by central differences that replay the
recorded ranks, i.e. four further evaluations per theta); value = log L + gradient evaluations per second
in the headline arithmetic (--precision, default plain FP64; FP32 and BF16 legs beside it: the
factorization's off-diagonal products and recompressions in FP32 or with bfloat16 product operands, each
pair guarded, POTRF / TRSM / diagonal updates / initial compression / log-det in FP64).

Also reported: the same step in FP32 and in BF16 (bfloat16 product operands, FP32 accumulation) on the
same theta, with their rates, the ratio to plain FP64, the largest |delta log L| and gradient differences
against plain FP64 and the PD agreement; the roofline of the dominant kernel group (one factorization step: the recompressions of
step k with the POTRF, inverse, TRSM and diagonal update of step k + 1, timed with HIP events inside the
timed region, algorithmic flops and bytes from the rank logs); the compression's Matern products; and a
CPU baseline (the NumPy/SciPy oracle timed on the host cores on one full evaluation at one theta).

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "adaptive-tlr-pilots_amd"))

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 157.3}   # MI355X dense FP64 / FP32 (AMD published)
HBM_PEAK_GBS = 8000.0
# HBM bytes per factorization step and theta (FETCH_SIZE + WRITE_SIZE of the factorization's dispatches, each
# counter calibrated against a device copy of known size, divided by steps and theta); measured with
# rocprofv3 --pmc on this workload in plain FP64
# (profiles/factorization_and_compression_n40000_b500_pmc_traffic.md).
STEP_TRAFFIC_PER_THETA = 0.435e9
# central-difference step of the gradient in log theta: 1e-5 in FP64; the reduced-precision modes take a
# larger step so that their rounding (not smooth in theta) stays small against the difference
GRAD_H = {"fp64": 1e-5, "fp32": 1e-3, "bf16": 1e-3}


def pd_sweep(ctx, m, batch, rule, value, seed=1):
    """m theta drawn uniformly from the box (sigma2 in [0.5, 2], range in [0.1, 0.6], eta in [0.01, 0.1]): for
    every off-diagonal arithmetic the PD rate, the log L error against the dense FP64 log L and |dlog L| against
    plain FP64 TLR; and the dense log L's own PD rate."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array([0.5, 0.1, 0.01]), np.array([2.0, 0.6, 0.1])
    th = lo + rng.random((m, 3)) * (hi - lo)
    dense = np.empty(m)
    for i0 in range(0, m, 20):
        d = ctx.dense_loglik_batch(th[i0:i0 + 20])
        dense[i0:i0 + 20] = d["loglik"]
        print(f"# pd sweep: dense log L {min(i0 + 20, m)}/{m}", file=sys.stderr, flush=True)
    res = {}
    for prec in ("fp64", "fp32", "bf16"):
        ll = np.empty(m)
        ok = np.empty(m, bool)
        for i0 in range(0, m, batch):
            r = ctx.loglik_batch(th[i0:i0 + batch], rule, value, precision=prec, ranks=False)
            ll[i0:i0 + batch], ok[i0:i0 + batch] = r["loglik"], r["completed"]
        print(f"# pd sweep: {prec} done", file=sys.stderr, flush=True)
        res[prec] = (ll, ok)
    ll64, ok64 = res["fp64"]
    dok = np.isfinite(dense)
    out = {"theta": m, "seed": seed, "box": {"sigma2": [0.5, 2.0], "range": [0.1, 0.6], "eta": [0.01, 0.1]},
           "dense_pd_rate": float(dok.mean()), "per_precision": {}}
    for prec, (ll, ok) in res.items():
        both = ok & dok
        err = np.abs(ll[both] - dense[both])
        e = {"pd_rate": float(ok.mean()), "pd_agree_with_fp64": bool(np.array_equal(ok, ok64)),
             "max_abs_err_vs_dense": float(err.max()) if err.size else None,
             "median_abs_err_vs_dense": float(np.median(err)) if err.size else None}
        if prec != "fp64":
            b2 = ok & ok64
            e["max_abs_dlogL_vs_fp64"] = float(np.abs(ll[b2] - ll64[b2]).max()) if b2.any() else None
        out["per_precision"][prec] = e
    return out


def lhs_thetas(m, seed=0):
    from scipy.stats import qmc
    u = qmc.LatinHypercube(d=3, seed=seed).random(m)
    lo = np.array([0.5, 0.1, 0.01])
    hi = np.array([2.0, 0.6, 0.1])
    return lo + u * (hi - lo)


def load_field(n):
    """Seed-0 synthetic field (S ~ U[0,1]^2 from default_rng(0), Morton order, y ~ N(0, Sigma(1, 0.3, 0.05))
    drawn by a dense FP64 Cholesky); stored as a fixture because the dense draw takes minutes on a CPU."""
    path = os.path.join(ROOT, "tests", "golden", f"seed0_n{n}.npy")
    a = np.load(path)
    return a[:, :2].copy(), a[:, 2].copy()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(S, y, b, theta, rule, value, n_sample, evals_per_point):
    """Time the CPU oracle (NumPy/SciPy restatement of the reference TLR Cholesky, all host cores for BLAS)
    on one full log L evaluation at theta (n_sample = N), or on the first n_sample sites of the field (Morton
    order keeps them a compact sub-region) scaled to N by the algorithmic work (tile count for the
    compression, C(nB, 3) for the recompressions); a log L + gradient costs evals_per_point evaluations (the
    same frozen-rank difference scheme)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tlr_oracle as O
    from threadpoolctl import threadpool_limits
    cores = os.cpu_count()
    n_full = len(y)
    t_start = time.perf_counter()

    def progress(phase, done, total):   # the oracle's own progress, every tenth tile row or step
        if done % 10 == 0 or done == total:
            print(f"# cpu baseline: {phase} {done}/{total} ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr,
                  flush=True)

    with threadpool_limits(limits=cores):
        Ss, ys = S[:n_sample], y[:n_sample]
        bnd = O.tile_bounds(n_sample, b)
        t0 = time.perf_counter()
        ll, _, info = O.tlr_loglik(Ss, ys, theta, bnd, rule, value, progress=progress)
        t = time.perf_counter() - t0
    th = [float(x) for x in theta]
    if n_sample >= n_full:
        return dict(value=1.0 / (evals_per_point * t), unit="evals/s", cores=cores, kind="port", cpu=_cpu_model(),
                    sample=(f"oracle tlr_loglik, one full evaluation at N={n_full}, b={b}, {rule} {value}, "
                            f"theta={th}: {t:.1f} s (log L {ll}); x{evals_per_point} log L evaluations per "
                            f"log L + gradient (the same frozen-rank difference scheme)"))
    # smaller sample: split compression and factorization, scale each by its work
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        sig2, r, eta = theta
        thr_rel, thr_abs = O.rule_thresholds(rule, value, theta, n_sample, b)
        for i, (i0, i1) in enumerate(bnd):
            for j, (j0, j1) in enumerate(bnd[:i]):
                T = sig2 * O.matern32(O.dist(Ss[i0:i1], Ss[j0:j1]), r)
                O.truncate_svd(T, rule, thr_rel, thr_abs)
        tc = time.perf_counter() - t0
    nbs, nbf = len(bnd), (n_full + b - 1) // b
    tiles_s, tiles_f = nbs * (nbs - 1) / 2, nbf * (nbf - 1) / 2
    rec_s, rec_f = nbs * (nbs - 1) * (nbs - 2) / 6, nbf * (nbf - 1) * (nbf - 2) / 6
    t_full = tc * tiles_f / tiles_s + max(t - tc, 0.0) * rec_f / max(rec_s, 1.0)
    return dict(value=1.0 / (evals_per_point * t_full), unit="evals/s", cores=cores, kind="port", cpu=_cpu_model(),
                sample=(f"oracle tlr_loglik at N={n_sample}, b={b}, {rule} {value}, theta={th}: {t:.1f} s "
                        f"({tc:.1f} s compression), scaled to N={n_full} by tile count and C(nB,3): "
                        f"{t_full:.0f} s per log L, x{evals_per_point} per log L + gradient"))


def run_leg(ctx, thetas, prec, rule, value, steps, warmup, grad, timing=False, tag=""):
    """warmup + steps batched evaluations in arithmetic prec; returns (seconds per step, last result,
    timing of the factorization steps when timing)."""
    h = GRAD_H[prec]

    def step():
        if grad:
            return ctx.loglik_grad_batch(thetas, rule, value, h=h, precision=prec)
        return ctx.loglik_batch(thetas, rule, value, precision=prec)

    for w in range(warmup):
        r = step()
        print(f"# {tag}{prec} warmup {w}: completed {int(r['completed'].sum())}/{len(thetas)}", file=sys.stderr,
              flush=True)
    if timing:
        ctx.set_timing(5)   # the factorization steps: the dominant kernel group
    t0 = time.perf_counter()
    for s in range(steps):
        r = step()
        print(f"# {tag}{prec} step {s}: {time.perf_counter() - t0:.3f} s", file=sys.stderr, flush=True)
    elapsed = (time.perf_counter() - t0) / steps
    tm = ctx.get_timing() if timing else None
    if timing:
        ctx.set_timing(0)
    return elapsed, r, tm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=40000)
    ap.add_argument("--b", type=int, default=500)
    ap.add_argument("--batch", type=int, default=20)
    ap.add_argument("--rule", default="adaptive")
    ap.add_argument("--value", type=float, default=0.9)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16"],
                    help="arithmetic of the headline leg (plain FP64: at this budget the guard lets FP32 take "
                         "only a few per cent of the recompressions, so the reduced modes are not faster)")
    ap.add_argument("--compare", default="fp32,bf16", help="comparison legs on the same theta ('' for none)")
    ap.add_argument("--compare-steps", type=int, default=1, help="timed steps of each comparison leg")
    ap.add_argument("--compare-warmup", type=int, default=0,
                    help="untimed steps of each comparison leg (the headline leg has loaded every kernel)")
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="sites of the CPU-baseline sample (-1: the full N, one evaluation; 0: skip)")
    ap.add_argument("--no-grad", action="store_true", help="time log L alone")
    ap.add_argument("--pd-sweep", type=int, default=0,
                    help="theta drawn uniformly from the box for a PD-rate and error sweep per arithmetic against "
                         "the dense FP64 log L (0: skip; 200 takes about 4 minutes at N = 40,000)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not supported (one GPU's share of the path is "
              "what this library runs)", file=sys.stderr)
        sys.exit(2)
    import tlrgp
    S, y = load_field(a.n)
    thetas = lhs_thetas(a.batch)
    print(f"# device: {tlrgp.device_info()}", file=sys.stderr, flush=True)
    grad = not a.no_grad
    # the gradient's 4 replays per theta run as one batch when the context holds 5 x batch slots
    # (about 1.8 GB per slot at N = 40,000, b = 500: 180 GB of the 288 GB for 20 theta)
    ctx = tlrgp.TLRContext(S, y, a.b, max_batch=a.batch * (5 if grad else 1))
    prec = a.precision
    t_step, r, (kms, launches, flops, kbytes) = run_leg(ctx, thetas, prec, a.rule, a.value, a.steps, a.warmup,
                                                        grad, timing=True)
    counts = [ctx.arith_counts(s) for s in range(a.batch)]
    legs = {}
    for cp in [c for c in a.compare.split(",") if c and c != prec]:
        tc, rc, _ = run_leg(ctx, thetas, cp, a.rule, a.value, a.compare_steps, a.compare_warmup, grad,
                            tag="compare ")
        legs[cp] = (tc, rc, [ctx.arith_counts(s) for s in range(a.batch)])
    # one more (untimed for throughput) pass timing the Matern products of the compression
    ctx.set_timing(1)
    ctx.loglik_batch(thetas, a.rule, a.value, precision=prec)
    pms, plaunches, pflops, pbytes = ctx.get_timing()
    ctx.set_timing(0)
    evals = a.batch / t_step
    nb = (a.n + a.b - 1) // a.b
    il = np.tril_indices(nb, -1)
    mean_rank = float(r["ranks_factor"][:, il[0], il[1]].mean())
    achieved_tf = flops / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    achieved_gbs = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    peak = PEAK_TFLOPS[prec]
    frac_f, frac_b = achieved_tf / peak, achieved_gbs / HBM_PEAK_GBS
    # theta-steps per timed launch group (steps with recompressions: k = 0 .. nB - 3)
    theta_steps = a.batch * (5 if grad else 1) * a.steps * max(nb - 2, 0)
    traffic = STEP_TRAFFIC_PER_THETA * theta_steps / max(launches, 1)
    if frac_b >= frac_f:
        roof = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": frac_b}
    else:
        roof = {"bound": "mfma", "achieved": achieved_tf, "peak": peak, "unit": "TFLOP/s", "frac": frac_f}
    roof.update({"traffic": traffic,
                 "kernel": "factorization step (work lists, k_recompress classes, lookahead POTRF/TRTRI/TRSM/"
                           "diagonal update)",
                 "launches": launches, "kernel_ms": kms, "flops_per_launch": flops / max(launches, 1),
                 "bytes_per_launch": kbytes / max(launches, 1), "achieved_tflops": achieved_tf,
                 "frac_of_peak_flops": frac_f, "achieved_gbs": achieved_gbs, "frac_of_hbm": frac_b,
                 "limiter": "latency of the dependent per-pair steps (neither roofline is reached)"})
    p_achieved = pflops / (pms * 1e-3) / 1e12 if pms > 0 else 0.0
    comparison = {}
    for cp, (tc, rc, cc) in legs.items():
        ok = r["completed"] & rc["completed"]
        entry = {"evals_per_s": a.batch / tc, "ms_per_step": 1e3 * tc,
                 "pd_agree": bool(np.array_equal(r["completed"], rc["completed"])),
                 "completed": int(rc["completed"].sum()),
                 "recompressions": {k: int(sum(c[k] for c in cc)) for k in ("fp64", "fp32", "bf16")}}
        if ok.any():
            entry["max_abs_dlogL_vs_headline"] = float(np.abs(r["loglik"][ok] - rc["loglik"][ok]).max())
            if grad:
                g1, g2 = r["grad"][ok], rc["grad"][ok]
                entry["max_rel_dgrad_vs_headline"] = float(
                    (np.abs(g1 - g2) / np.maximum(np.abs(g2), 1.0)).max())
        comparison[cp] = entry
    # mixed (reduced-precision) rate over plain FP64, whichever of them is the headline leg
    if prec == "fp64":
        mixed = comparison.get("fp32", {}).get("evals_per_s")
        ratio = (mixed / evals) if mixed else None
    else:
        fp64_rate = comparison.get("fp64", {}).get("evals_per_s")
        ratio = (evals / fp64_rate) if fp64_rate else None
    out = {
        "metric": "TLR logL+grad evals/s at N=40k, mixed vs plain FP64 GPU, |dlogL|<=1e-2",
        "value": evals, "unit": "evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * t_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16": "bf16+f32+f64"}[prec],
        "data": "synthetic (seed-0 uniform sites, Morton order, dense FP64 draw of y)",
        "config": {"workload": f"TLR log-likelihood{' + gradient' if grad else ''}, N={a.n}, b={a.b}, "
                               f"{a.rule} tau={a.value}, {a.batch} Latin-hypercube theta per step, "
                               f"off-diagonal arithmetic {prec}",
                   "n": a.n, "b": a.b, "theta_per_step": a.batch, "precision": prec, "gradient": grad,
                   "gradient_method": f"frozen ranks and arithmetic: sigma2 closed form, range/eta central "
                                      f"differences h={GRAD_H[prec]}" if grad else None,
                   "completed": int(r["completed"].sum()), "mean_rank_factor": mean_rank,
                   "recompressions": {k: int(sum(c[k] for c in counts)) for k in ("fp64", "fp32", "bf16")}},
        "mixed_vs_plain_fp64": ratio,
        "comparison": comparison,
        "roofline": roof,
        "roofline_compression": {"bound": "mfma", "achieved": p_achieved, "peak": PEAK_TFLOPS["fp64"],
                                 "unit": "TFLOP/s", "frac": p_achieved / PEAK_TFLOPS["fp64"],
                                 "kernel": "k_matern_product", "launches": plaunches, "kernel_ms": pms,
                                 "bytes_per_launch": pbytes / max(plaunches, 1),
                                 "achieved_gbs": pbytes / (pms * 1e-3) / 1e9 if pms > 0 else 0.0},
    }
    if a.pd_sweep > 0:
        out["pd_sweep"] = pd_sweep(ctx, a.pd_sweep, ctx.max_batch, a.rule, a.value)
    if a.cpu_sample:
        n_s = a.n if a.cpu_sample < 0 else a.cpu_sample
        out["cpu_baseline"] = cpu_baseline(S, y, a.b, tuple(thetas[0]), a.rule, a.value, n_s, 5 if grad else 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
