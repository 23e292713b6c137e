"""Write the seed-0 synthetic fields of the benchmark configurations (S ~ U[0,1]^2 from default_rng(0),
Morton order, y ~ N(0, Sigma(1, 0.3, 0.05)) by a dense FP64 Cholesky) as seed0_n{N}.npy (columns x, y, z).

The dense draw takes minutes at N = 40,000, so the fields are stored rather than regenerated by the
benchmark.  Usage:  python3 tests/golden/make_seed0.py [N ...]"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402


def main(ns):
    for n in ns:
        S, y = O.simulate_seed0(n, seed=0)
        np.save(os.path.join(HERE, f"seed0_n{n}.npy"), np.column_stack([S, y]))
        print("wrote", n)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [10000, 40000])
