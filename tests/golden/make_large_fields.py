"""The reference's two largest TLR-Cholesky runs (b=250, theta=(1, 0.3, 0.05), seed 20261020): the
Morton-ordered fields results/tlr_cholesky/sim_n{60000,100000}_obs.csv (text, %.18e: an exact float64 round
trip) as sim_n{N}_obs.npy, and their result JSONs' per-rule completion, stored values, mean/max factor rank,
log L and failing tile as the tlr_cholesky entries "60000" and "100000" of goldens.json (read as text, no
code is run).

Usage:  python3 tests/golden/make_large_fields.py [/root/reference]"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = ("rule", "value", "completed", "stored_sigma", "stored_factor", "mean_rank_factor", "max_rank_factor",
        "loglik", "loglik_err", "failed_at_tile")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(HERE, "goldens.json")
    with open(path) as f:
        g = json.load(f)
    for n in (60000, 100000):
        a = np.loadtxt(os.path.join(ref, "results", "tlr_cholesky", f"sim_n{n}_obs.csv"), delimiter=",", skiprows=1)
        np.save(os.path.join(HERE, f"sim_n{n}_obs.npy"), a)
        with open(os.path.join(ref, "results", "tlr_cholesky", f"tlr_cholesky_n{n}.json")) as f:
            r = json.load(f)
        g["tlr_cholesky"][str(n)] = dict(dense_loglik=r.get("dense_loglik_true"),
                                         rules=[{k: x[k] for k in KEYS if k in x} for x in r["rules"]])
    with open(path, "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
