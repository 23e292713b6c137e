"""Autotuner of a TLR context's tunables (no counterpart in the reference, which has no GPU path).

Three kinds of tunables (TLRContext.set_tunable, C ABI tlr_ctx_set_tunable):

* launch shape -- ``diag_wg`` (work groups per tile of the diagonal update) and ``grid_frac`` (persistent
  recompression grids as a fraction of the resident capacity).  They change which work group does what, not
  the arithmetic of any entry, so every setting gives the same log L bit for bit (checked here); the fastest
  is kept.
* accuracy -- ``guard``: the headroom (in unit roundoffs of the working precision) that a reduced-precision
  recompression needs, else it is redone at the next higher precision.  A smaller guard sends more pairs to
  FP32 and moves log L further from plain FP64.  The tuner evaluates the sample theta in FP64 once and takes
  the smallest guard whose largest |log L - log L_FP64| stays within margin * bound with the same PD flags.
* sketch width -- ``sketch``: columns of the first compression round.  A rank is accepted only with 24
  sketch columns to spare (else the tile goes to the second round at min(96, b) columns), so the width
  moves work between the rounds and changes the rounding of the compression, not its ranks.  The tuner
  keeps the fastest width whose initial and final ranks equal the default width's on every sample theta and
  whose log L agrees with it to a relative 1e-9.

Usage: python -m tlrgp.autotune --field tests/golden/seed0_n40000.npy --b 500 --batch 20 --out tuning.json
"""
import argparse
import json
import time

import numpy as np

GUARDS = (256.0, 512.0, 1024.0, 2048.0, 4096.0, 16384.0, 1e6)
DIAG_WGS = (3, 6, 9, 12, 18)
GRID_FRACS = (0.5, 0.75, 1.0)
SKETCHES = (64, 56, 48, 40, 32)


def _timed(ctx, thetas, rule, value, precision, reps):
    best, r = None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = ctx.loglik_batch(thetas, rule, value, precision, ranks=False)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, r


def tune_guard(ctx, thetas, rule="adaptive", value=0.9, precision="fp32", guards=GUARDS, bound=1e-2, margin=0.5,
               reps=1, log=print):
    """Smallest guard with max |delta log L| <= margin * bound against FP64 and identical PD flags on thetas.
    Leaves the chosen guard set on ctx; returns (guard, table)."""
    t64, ref = _timed(ctx, thetas, rule, value, "fp64", reps)
    log(f"autotune guard: fp64 {t64:.3f} s per batch")
    table = []
    chosen = None
    for g in sorted(guards):
        ctx.set_tunable("guard", g)
        dt, r = _timed(ctx, thetas, rule, value, precision, reps)
        ok = ref["completed"] & r["completed"]
        dl = float(np.abs(r["loglik"][ok] - ref["loglik"][ok]).max()) if ok.any() else 0.0
        pd = bool(np.array_equal(ref["completed"], r["completed"]))
        cnt = [ctx.arith_counts(s) for s in range(len(thetas))]
        frac = sum(c["fp32"] + c["bf16"] for c in cnt) / max(1, sum(sum(c.values()) for c in cnt))
        row = dict(guard=g, seconds=dt, fp64_seconds=t64, max_abs_dlogL=dl, pd_agree=pd, reduced_frac=frac)
        table.append(row)
        log(f"autotune guard {g:g}: {dt:.3f} s, max |dlogL| {dl:.3e}, PD agree {pd}, reduced-precision pairs "
            f"{100 * frac:.1f} %")
        if chosen is None and pd and dl <= margin * bound:
            chosen = g
    if chosen is None:
        chosen = max(guards)
    ctx.set_tunable("guard", chosen)
    return chosen, table


def tune_launch(ctx, thetas, rule="adaptive", value=0.9, precision="fp64", diag_wgs=DIAG_WGS, grid_fracs=GRID_FRACS,
                reps=1, log=print):
    """Fastest diag_wg, then fastest grid_frac (coordinate search); every setting must reproduce the first
    setting's log L bit for bit.  Leaves the chosen values set; returns (dict, table)."""
    table = []
    base = None
    best = {}
    for name, values in (("diag_wg", diag_wgs), ("grid_frac", grid_fracs)):
        times = []
        for v in values:
            ctx.set_tunable(name, v)
            dt, r = _timed(ctx, thetas, rule, value, precision, reps)
            if base is None:
                base = r["loglik"].copy()
            same = bool(np.array_equal(np.nan_to_num(base), np.nan_to_num(r["loglik"])))
            if not same:
                raise RuntimeError(f"tunable {name}={v} changed log L: the launch shape must not change results")
            times.append(dt)
            table.append(dict(name=name, value=v, seconds=dt))
            log(f"autotune {name} {v}: {dt:.3f} s")
        best[name] = values[int(np.argmin(times))]
        ctx.set_tunable(name, best[name])
    return best, table


def tune_sketch(ctx, thetas, rule="adaptive", value=0.9, precision="fp64", widths=SKETCHES, rtol=1e-9, reps=1,
                log=print):
    """Fastest first-round sketch width (the first of widths is the reference) whose initial and final ranks
    equal the reference's on thetas, with the same PD flags and log L within rtol relative.  Leaves the
    chosen width set; returns (width, table)."""
    table, ref, times = [], None, []
    for k in widths:
        ctx.set_tunable("sketch", k)
        best, r = None, None
        for _ in range(reps):
            t0 = time.perf_counter()
            r = ctx.loglik_batch(thetas, rule, value, precision)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        if ref is None:
            ref = r
        ok = ref["completed"] & r["completed"]
        same_ranks = bool(np.array_equal(ref["ranks_sigma"], r["ranks_sigma"]) and
                          np.array_equal(ref["ranks_factor"][ok], r["ranks_factor"][ok]))
        pd = bool(np.array_equal(ref["completed"], r["completed"]))
        rel = float((np.abs(r["loglik"][ok] - ref["loglik"][ok]) / np.maximum(np.abs(ref["loglik"][ok]), 1.0)).max()) \
            if ok.any() else 0.0
        accepted = same_ranks and pd and rel <= rtol
        table.append(dict(sketch=k, seconds=best, same_ranks=same_ranks, pd_agree=pd, max_rel_dlogL=rel,
                          accepted=accepted))
        times.append(best if accepted else np.inf)
        log(f"autotune sketch {k}: {best:.3f} s, ranks equal {same_ranks}, PD agree {pd}, max rel dlogL {rel:.2e}")
    chosen = widths[int(np.argmin(times))]
    ctx.set_tunable("sketch", chosen)
    return chosen, table


def apply(ctx, tuning):
    """Set the tunables of a tuning dict (as written by main) on ctx."""
    for name in ("guard", "diag_wg", "grid_frac", "sketch"):
        if name in tuning.get("chosen", {}):
            ctx.set_tunable(name, tuning["chosen"][name])


def main(argv=None):
    from .likelihood import TLRContext
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--field", required=True, help=".npy with columns x, y, z (Morton-ordered sites, observations)")
    ap.add_argument("--b", type=int, default=500)
    ap.add_argument("--batch", type=int, default=20)
    ap.add_argument("--rule", default="adaptive")
    ap.add_argument("--value", type=float, default=0.9)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--bound", type=float, default=1e-2)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1, help="Latin-hypercube seed of the tuning theta")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    from scipy.stats import qmc
    f = np.load(a.field)
    S, y = f[:, :2].copy(), f[:, 2].copy()
    u = qmc.LatinHypercube(d=3, seed=a.seed).random(a.batch)
    thetas = np.array([0.5, 0.1, 0.01]) + u * np.array([1.5, 0.5, 0.09])
    ctx = TLRContext(S, y, a.b, max_batch=a.batch)
    launch, t1 = tune_launch(ctx, thetas, a.rule, a.value, "fp64", reps=a.reps, log=lambda m: print(m, flush=True))
    sketch, t3 = tune_sketch(ctx, thetas, a.rule, a.value, "fp64", reps=a.reps, log=lambda m: print(m, flush=True))
    guard, t2 = tune_guard(ctx, thetas, a.rule, a.value, a.precision, bound=a.bound, reps=a.reps,
                           log=lambda m: print(m, flush=True))
    out = dict(n=len(y), b=a.b, batch=a.batch, rule=a.rule, value=a.value, precision=a.precision, bound=a.bound,
               chosen=dict(launch, guard=guard, sketch=sketch), launch_table=t1, guard_table=t2, sketch_table=t3)
    print(json.dumps(out), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(out, fh, indent=1)
    return out


if __name__ == "__main__":
    main()
