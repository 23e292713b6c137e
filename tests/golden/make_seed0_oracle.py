"""Oracle results on the seed-0 N=40,000 field at the benchmark tile size b=500 (adaptive tau=0.9,
theta=(1, 0.3, 0.05)): log L, log-det, quadratic form and the per-tile ranks (initial and final), written
to seed0_n40000_b500_oracle.npz for the full-size GPU parity test.  Takes a few minutes on 8 cores.

Usage:  python3 tests/golden/make_seed0_oracle.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402


def main():
    a = np.load(os.path.join(HERE, "seed0_n40000.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    theta = np.array([1.0, 0.3, 0.05])
    ll, _, info = O.tlr_loglik(S, y, theta, O.tile_bounds(40000, 500), "adaptive", 0.9)
    np.savez_compressed(os.path.join(HERE, "seed0_n40000_b500_oracle.npz"), theta=theta, loglik=ll,
                        logdet=info["logdet"], quad=info["quad"],
                        ranks_sigma=info["ranks_sigma"].astype(np.int16),
                        ranks_factor=info["ranks_factor"].astype(np.int16))
    print("loglik", ll, "mean rank", info["mean_rank_factor"])


if __name__ == "__main__":
    main()
