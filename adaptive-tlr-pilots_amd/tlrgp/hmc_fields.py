"""Batched HMC over chains on several fields at once (no counterpart in the reference, which runs one
Metropolis-Hastings chain per process: codes/38_tlr_cholesky_mcmc.py:89-126, codes/19_tlr_mcmc_gpu.py:170-203).

The chains of each field share one TLRContext; every leapfrog step evaluates log L and its gradient for all
chains of all fields, each field's batch on its own context at the same time (inference.concurrent_loglik_grad).
Reports evaluations per second and the p50 / p99 latency of a leapfrog step.

Usage (fields as .npy with columns x, y, z; chains:tile-size per field):
  python -m tlrgp.hmc_fields --field a.npy:64:250 --field b.npy:16:500 --leapfrog 10 --iters 1
"""
import argparse
import json
import time

import numpy as np

from . import inference
from .likelihood import TLRContext


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--field", action="append", required=True, help="path.npy:chains:b")
    ap.add_argument("--iters", type=int, default=1, help="HMC iterations")
    ap.add_argument("--leapfrog", type=int, default=10, help="leapfrog steps per iteration")
    ap.add_argument("--step", type=float, default=0.01, help="leapfrog step in log theta")
    ap.add_argument("--rule", default="adaptive")
    ap.add_argument("--value", type=float, default=0.9)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--replay-chunks", type=int, default=2,
                    help="a field's context holds chains x (1 + 4 / chunks) slots: the gradient's 4 replays per "
                         "chain run in this many batches")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ckpt", default=None, help="state file: saved every --ckpt-every iterations, resumed from")
    ap.add_argument("--ckpt-every", type=int, default=1)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    lo, hi = np.array([0.5, 0.1, 0.01]), np.array([2.0, 0.6, 0.1])
    parts, starts, names = [], [], []
    for spec in a.field:
        path, k, b = spec.split(":")
        k, b = int(k), int(b)
        f = np.load(path)
        ctx = TLRContext(f[:, :2].copy(), f[:, 2].copy(), b, max_batch=max(k, (4 * k) // a.replay_chunks))
        h = 1e-5 if a.precision == "fp64" else 1e-3
        parts.append((inference.batched_loglik_grad(ctx, a.rule, a.value, h=h, precision=a.precision), k))
        starts.append(np.log(lo + rng.random((k, 3)) * (hi - lo)))
        names.append(f"{k} chains x N={len(f)} (b={b})")
        print(f"# field {path}: {names[-1]}", flush=True)
    f = inference.concurrent_loglik_grad(parts)
    t0 = time.perf_counter()
    res = inference.hmc(f, np.concatenate(starts), a.iters, step=a.step, n_leapfrog=a.leapfrog, seed=a.seed,
                        ckpt=a.ckpt, ckpt_every=a.ckpt_every,
                        log=lambda m: print(f"# {m} ({time.perf_counter() - t0:.0f} s)", flush=True))
    wall = time.perf_counter() - t0
    out = dict(fields=names, precision=a.precision, iters=a.iters, leapfrog=a.leapfrog, step=a.step,
               evaluations=res["evaluations"], evals_per_s=res["evaluations"] / res["seconds"],
               step_latency_p50=res["step_latency_p50"], step_latency_p99=res["step_latency_p99"],
               wall_seconds=wall, acceptance_mean=float(np.mean(res["acceptance"])),
               rank_capped_evaluations=int(sum(getattr(fn, "rank_capped", 0) for fn, _ in parts)))
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
