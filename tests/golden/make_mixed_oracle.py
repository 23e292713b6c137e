"""Oracle fixture of the storage-precision rule mixed:0.9 on the reference's N = 10k field (b = 250,
theta = (1, 0.3, 0.05)): log L, log-det, quadratic form, initial and final ranks and the storage precision
of every factor tile, written to mixed_n10000_oracle.npz.  The GPU test compares the library with it tile
for tile (the oracle run takes about 40 s on 8 cores, too long for the GPU test itself).

Usage:  python3 tests/golden/make_mixed_oracle.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402


def main():
    a = np.load(os.path.join(HERE, "sim_n10000_obs.npy"))
    S, y = a[:, :2], a[:, 2]
    theta = np.array([1.0, 0.3, 0.05])
    ll, _, info = O.tlr_loglik(S, y, theta, O.tile_bounds(10000, 250), "mixed", 0.9)
    np.savez(os.path.join(HERE, "mixed_n10000_oracle.npz"), theta=theta, loglik=ll, logdet=info["logdet"],
             quad=info["quad"], ranks_sigma=info["ranks_sigma"].astype(np.int16),
             ranks_factor=info["ranks_factor"].astype(np.int16),
             precision_factor=info["precision_factor"].astype(np.int16),
             stored_factor_bytes=info["stored_factor_bytes"])
    print("log L", ll, "mix", info["precision_mix_factor_tiles"])


if __name__ == "__main__":
    main()
