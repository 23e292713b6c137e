"""Oracle results for three of the benchmark's Latin-hypercube theta on the seed-0 N=40,000 field at b=500
(adaptive tau=0.9): theta, log L, log-det, quadratic form and the per-tile ranks (initial and final), written
to seed0_n40000_b500_lhs_oracle.npz for the GPU parity test of the benchmark's own theta.  The theta are
rows 0, 7 and 13 of scipy.stats.qmc.LatinHypercube(d=3, seed=0).random(20) scaled to sigma2 in [0.5, 2],
range in [0.1, 0.6], eta in [0.01, 0.1] (the benchmark's sample).  Takes about 5 minutes per theta on 8 cores.

Usage:  python3 tests/golden/make_seed0_lhs_oracle.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402

ROWS = (0, 7, 13)


def lhs_thetas(m, seed=0):
    from scipy.stats import qmc
    u = qmc.LatinHypercube(d=3, seed=seed).random(m)
    return np.array([0.5, 0.1, 0.01]) + u * np.array([1.5, 0.5, 0.09])


def main():
    a = np.load(os.path.join(HERE, "seed0_n40000.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    thetas = lhs_thetas(20)[list(ROWS)]
    res = dict(theta=thetas, rows=np.array(ROWS), loglik=[], logdet=[], quad=[], ranks_sigma=[], ranks_factor=[])
    for th in thetas:
        ll, _, info = O.tlr_loglik(S, y, th, O.tile_bounds(40000, 500), "adaptive", 0.9)
        res["loglik"].append(ll)
        res["logdet"].append(info["logdet"])
        res["quad"].append(info["quad"])
        res["ranks_sigma"].append(info["ranks_sigma"].astype(np.int16))
        res["ranks_factor"].append(info["ranks_factor"].astype(np.int16))
        print("theta", th, "loglik", ll, "mean rank", info["mean_rank_factor"], flush=True)
    np.savez_compressed(os.path.join(HERE, "seed0_n40000_b500_lhs_oracle.npz"),
                        **{k: np.asarray(v) for k, v in res.items()})


if __name__ == "__main__":
    main()
