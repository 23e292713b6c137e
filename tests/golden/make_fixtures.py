"""Extract the golden fixtures of the TLR likelihood path from the reference's committed results.

Reads (text only, no code is run) the reference's ``results/`` folder:

* ``tlr_cholesky/sim_n{N}_obs.csv`` (N = 10k, 20k, 40k) and ``smoke/sim_n2000_obs.csv``: Morton-ordered
  sites and observations written with ``%.18e`` (an exact float64 round trip), stored here as
  ``sim_n{N}_obs.npy`` (float64, columns x, y, z);
* the paired result JSONs: dense and TLR log-likelihoods, stored values, mean/max factor ranks,
  completion flags and failing tiles per rule, the TLR log-determinant of ``kl_tlr.json``, the dense
  quadratic form of ``bound_terms_n{N}.json``.  Written to ``goldens.json``.

Usage:  python3 tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _load_json(ref, rel):
    with open(os.path.join(ref, "results", rel)) as f:
        return json.load(f)


def main(ref="/root/reference"):
    for n, rel in ((2000, "smoke/sim_n2000_obs.csv"), (10000, "tlr_cholesky/sim_n10000_obs.csv"),
                   (20000, "tlr_cholesky/sim_n20000_obs.csv"), (40000, "tlr_cholesky/sim_n40000_obs.csv")):
        a = np.loadtxt(os.path.join(ref, "results", rel), delimiter=",", skiprows=1)
        assert a.shape == (n, 3), a.shape
        np.save(os.path.join(HERE, f"sim_n{n}_obs.npy"), a)

    g = dict(source="reference results/ (seed 20261020, theta=(1, 0.3, 0.05), Matern nu=3/2, b=250)", tlr_cholesky={})
    keep = ("rule", "value", "completed", "stored_sigma", "stored_factor", "mean_rank_factor", "max_rank_factor",
            "loglik", "loglik_err", "failed_at_tile")
    for n in (10000, 20000, 40000):
        d = _load_json(ref, f"tlr_cholesky/tlr_cholesky_n{n}.json")
        rules = [{k: r[k] for k in keep if k in r} for r in d["rules"]]
        if n == 40000:   # the eps rerun adds fixed 1e-5 / 1e-6 at N=40k
            e = _load_json(ref, "tlr_cholesky_eps/tlr_cholesky_n40000.json")
            rules += [{k: r[k] for k in keep if k in r} for r in e["rules"] if r["rule"] == "fixed"]
        g["tlr_cholesky"][str(n)] = dict(b=d["settings"]["b"], dense_loglik=d["dense_loglik_true"], rules=rules)
    kl = _load_json(ref, "screening_stress/nu15_r03_e005/kl_tlr.json")
    g["kl_tlr_n10000"] = dict(logdet_exact=kl["logdet_exact"], logdet_approx=kl["logdet_approx"],
                              tau=kl["tau"], note="screening field nu15_r03_e005 = the same N=10k field")
    g["bound_terms"] = {}
    for n in (10000, 40000):
        bt = _load_json(ref, f"bound_terms/bound_terms_n{n}.json")
        g["bound_terms"][str(n)] = dict(dense_loglik=bt["dense_loglik"], z2=bt["z2"])
    sm = _load_json(ref, "smoke/tlr_rank_rules_n2000.json")
    g["smoke_rank_rules_n2000"] = [dict(rule=r["rule"], value=r["value"], mean_rank=r["mean_rank"],
                                        max_rank=r["max_rank"], min_rank=r["min_rank"], pd=r["pd"],
                                        max_tile_abs_err=r["max_tile_abs_err"], loglik_err=r["loglik_err"])
                                   for r in sm["rules"]]
    with open(os.path.join(HERE, "goldens.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(*sys.argv[1:])
