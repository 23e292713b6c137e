"""Extract the golden fixtures of the TLR likelihood path from the reference's committed results.

Reads (text only, no code is run) the reference's ``results/`` folder:

* ``tlr_cholesky/sim_n{N}_obs.csv`` (N = 10k, 20k, 40k) and ``smoke/sim_n2000_obs.csv``: Morton-ordered
  sites and observations written with ``%.18e`` (an exact float64 round trip), stored here as
  ``sim_n{N}_obs.npy`` (float64, columns x, y, z); the 300 held-out points of the N = 10k and 40k runs
  as ``sim_n{N}_holdout.npy``;
* the paired result JSONs: dense and TLR log-likelihoods, stored values, mean/max factor ranks,
  completion flags and failing tiles per rule, the TLR log-determinant of ``kl_tlr.json`` (the N = 10k
  field and the three other Matern-3/2 fields of ``screening_stress/``, which share its sites), the dense
  quadratic form of ``bound_terms_n{N}.json`` and that file's error terms of every rule (N = 10k, 40k)
  (codes/36_likelihood_error_terms.py), the Nelder-Mead MLE results (theta_hat, objective,
  evaluation and failure counts) of the N = 10k and 20k runs, the ``mixed`` storage-precision rule's
  results at N = 10k and 20k (``tlr_cholesky_mixed/``), and the initial-rank aggregates (mean and max
  rank of the adaptive rule at 0.9 and the fixed rule at 1e-4) of every Matern-3/2 run of ``joint_sweep/``
  (N = 10k, the same sites, b = 250 and 500, ranges 0.05-3, nuggets 0.05-1) whose ranks stay within 64,
  and for the four Matern-3/2 fields of ``screening_stress/`` the TLR maximum-likelihood point of
  ``inference.json`` (theta_hat with the adaptive rule's log L and the dense log L there; the fields'
  observations, on the N = 10k sites, as ``screening_<tag>_z.npy``).  Written to ``goldens.json``;
* ``mcmc_tlrchol/<rule>/c<k>/``: the first 1200 draws of the reference's random-walk chains at N = 5000
  (adaptive 0.9 chains 1-4, fixed 1e-4 chains 1 and 4) with their settings, as ``mcmc_n5000_heads.npz``
  and the ``mcmc_n5000`` entry of ``goldens.json``.

The N = 5000 field of those chains is not stored in the reference; it is regenerated with the oracle's
restatement of the reference's generator (seed 20261020, 300 held-out points) as ``sim_n5000_obs.npy`` /
``sim_n5000_holdout.npy``.  The chains themselves pin it: a different field would change their accept
decisions.

Usage:  python3 tests/golden/make_fixtures.py [/root/reference], then make_large_fields.py (which adds the
N = 60k and 100k entries to goldens.json).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEAD = 1200   # draws kept per reference chain: covers the Robbins-Monro phase and the re-shaping at 1000


def _load_json(ref, rel):
    with open(os.path.join(ref, "results", rel)) as f:
        return json.load(f)


def main(ref="/root/reference"):
    for n, rel in ((2000, "smoke/sim_n2000_obs.csv"), (10000, "tlr_cholesky/sim_n10000_obs.csv"),
                   (20000, "tlr_cholesky/sim_n20000_obs.csv"), (40000, "tlr_cholesky/sim_n40000_obs.csv")):
        a = np.loadtxt(os.path.join(ref, "results", rel), delimiter=",", skiprows=1)
        assert a.shape == (n, 3), a.shape
        np.save(os.path.join(HERE, f"sim_n{n}_obs.npy"), a)

    for n in (10000, 40000):   # held-out points of the same runs (300 each)
        a = np.loadtxt(os.path.join(ref, "results", f"tlr_cholesky/sim_n{n}_holdout.csv"), delimiter=",", skiprows=1)
        np.save(os.path.join(HERE, f"sim_n{n}_holdout.npy"), a)

    g = dict(source="reference results/ (seed 20261020, theta=(1, 0.3, 0.05), Matern nu=3/2, b=250)", tlr_cholesky={})
    keep = ("rule", "value", "completed", "stored_sigma", "stored_factor", "mean_rank_factor", "max_rank_factor",
            "loglik", "loglik_err", "failed_at_tile", "skill_true_theta")
    for n in (10000, 20000, 40000):
        d = _load_json(ref, f"tlr_cholesky/tlr_cholesky_n{n}.json")
        rules = [{k: r[k] for k in keep if k in r} for r in d["rules"]]
        if n == 40000:   # the eps rerun adds fixed 1e-5 / 1e-6 at N=40k
            e = _load_json(ref, "tlr_cholesky_eps/tlr_cholesky_n40000.json")
            rules += [{k: r[k] for k in keep if k in r} for r in e["rules"] if r["rule"] == "fixed"]
        g["tlr_cholesky"][str(n)] = dict(b=d["settings"]["b"], dense_loglik=d["dense_loglik_true"], rules=rules,
                                         dense_skill=d.get("dense_skill_true_theta"))
    kl = _load_json(ref, "screening_stress/nu15_r03_e005/kl_tlr.json")
    g["kl_tlr_n10000"] = dict(logdet_exact=kl["logdet_exact"], logdet_approx=kl["logdet_approx"], kl=kl["kl"],
                              trace=kl["trace"], tau=kl["tau"],
                              note="screening field nu15_r03_e005 = the same N=10k field")
    # the other Matern-3/2 screening fields: the KL certificate depends on the sites and theta only, and their
    # sites are those of the N = 10k field
    base = np.loadtxt(os.path.join(ref, "results", "tlr_cholesky", "sim_n10000_obs.csv"), delimiter=",",
                      skiprows=1)[:, :2]
    g["kl_tlr_screening"] = {}
    for tag in ("nu15_r03_e05", "nu15_r1_e005", "nu15_r3_e005"):
        st = _load_json(ref, f"screening_stress/{tag}/stress.json")["settings"]
        sites = np.loadtxt(os.path.join(ref, "results", "screening_stress", tag, "sim_n10000_obs.csv"),
                           delimiter=",", skiprows=1)[:, :2]
        assert st["nu"] == 1.5 and st["n"] == 10000 and np.array_equal(sites, base), tag
        kl = _load_json(ref, f"screening_stress/{tag}/kl_tlr.json")
        g["kl_tlr_screening"][tag] = dict(theta=[st["sigma2"], st["range"], st["eta"]], b=st["b"],
                                          logdet_exact=kl["logdet_exact"], logdet_approx=kl["logdet_approx"],
                                          kl=kl["kl"], trace=kl["trace"], tau=kl["tau"])
    # the TLR MLE points of the same four fields (codes/27_inference_headtohead.py: tlr_loglik of 17 with the
    # adaptive rule, and the dense log L, at the TLR theta_hat), with each field's observations
    g["screening_tlr_mle"] = {}
    for tag in ("nu15_r03_e005", "nu15_r03_e05", "nu15_r1_e005", "nu15_r3_e005"):
        st = _load_json(ref, f"screening_stress/{tag}/stress.json")["settings"]
        a = np.loadtxt(os.path.join(ref, "results", "screening_stress", tag, "sim_n10000_obs.csv"), delimiter=",",
                       skiprows=1)
        assert np.array_equal(a[:, :2], base), tag
        np.save(os.path.join(HERE, f"screening_{tag}_z.npy"), a[:, 2].copy())
        inf = _load_json(ref, f"screening_stress/{tag}/inference.json")
        t = inf["tlr_mle"]
        g["screening_tlr_mle"][tag] = dict(b=st["b"], tau=inf["tau"], theta_hat=t["theta_hat"],
                                           tlr_loglik=t["tlr_loglik"], exact_loglik=t["exact_loglik"])
    # initial ranks of the joint rank/precision sweep (codes/18_joint_rank_precision_pilot.py): rank_only is
    # the adaptive rule's truncation, fixed the fixed rule's, both by full SVD of every off-diagonal tile
    g["joint_sweep_ranks"] = []
    seen = set()
    for d in sorted(os.listdir(os.path.join(ref, "results", "joint_sweep"))):
        j = _load_json(ref, f"joint_sweep/{d}/joint_rank_precision_n10000.json")
        st = j["settings"]
        key = (st["b"], st["range"], st["eta"])
        if st.get("nu", 1.5) != 1.5 or key in seen:
            continue
        seen.add(key)
        assert st["n"] == 10000 and st["seed"] == 20261020 and st["sigma2"] == 1.0, d
        ra, rf = j["rules"]["rank_only"], j["rules"]["fixed"]
        if max(ra["max_rank"], rf["max_rank"]) > 64:   # beyond the 64-column rank capacity of the GPU slots
            continue
        g["joint_sweep_ranks"].append(dict(run=d, b=st["b"], theta=[st["sigma2"], st["range"], st["eta"]],
                                           tau=st["tau"], eps=st["eps"],
                                           adaptive=dict(mean_rank=ra["mean_rank"], max_rank=ra["max_rank"]),
                                           fixed=dict(mean_rank=rf["mean_rank"], max_rank=rf["max_rank"])))
    g["bound_terms"] = {}
    for n in (10000, 40000):
        bt = _load_json(ref, f"bound_terms/bound_terms_n{n}.json")
        # the dense terms and the error terms of every rule (codes/36_likelihood_error_terms.py)
        g["bound_terms"][str(n)] = dict(
            dense_loglik=bt["dense_loglik"], z2=bt["z2"], theta=bt["theta"], b=bt["settings"]["b"],
            sum_Kinv_tile_norms=bt["sum_Kinv_tile_norms"], sum_alpha_block_products=bt["sum_alpha_block_products"],
            rules={k: {kk: vv for kk, vv in v.items() if kk != "seconds"} for k, v in bt["rules"].items()})
    g["mixed"] = {}   # the storage-precision rule mixed:0.9 (adaptive ranks + FP8/16/32/64 factor storage)
    for n in (10000, 20000):
        d = _load_json(ref, f"tlr_cholesky_mixed/rules_only/tlr_cholesky_n{n}.json") if n == 10000 else \
            _load_json(ref, f"tlr_cholesky_mixed/tlr_cholesky_n{n}.json")
        r = next(x for x in d["rules"] if x["rule"] == "mixed")
        g["mixed"][str(n)] = {k: r[k] for k in ("value", "completed", "stored_sigma", "stored_factor",
                                               "stored_factor_bytes", "precision_mix_factor_tiles",
                                               "mean_rank_factor", "max_rank_factor", "loglik", "loglik_err")}
    g["mle"] = {}
    for n in (10000, 20000):
        d = _load_json(ref, f"tlr_cholesky/tlr_cholesky_n{n}.json")
        g["mle"][str(n)] = dict(x0_mult=[1.5, 0.6, 2.0], theta_true=d["theta_true"],
                                fits=[{k: f[k] for k in ("rule", "value", "theta_hat", "neg_loglik", "evals",
                                                         "failures", "skill") if k in f} for f in d["mle"]])
    chains, meta = {}, []
    for rd, cs in (("adaptive_0.9", (1, 2, 3, 4)), ("fixed_1e-4", (1, 4))):
        for c in cs:
            m = _load_json(ref, f"mcmc_tlrchol/{rd}/c{c}/tlr_mcmc_n5000.json")
            key = f"{rd}_c{c}"
            chains[key] = np.load(os.path.join(ref, "results", f"mcmc_tlrchol/{rd}/c{c}/chain_n5000.npy"))[:HEAD]
            st = m["settings"]
            meta.append(dict(key=key, rule=m["rule"], value=m["value"], chain_seed=st["chain_seed"],
                             start_theta=m["start_theta"], start_mult=[float(v) for v in st["start_mult"].split(",")],
                             iters=st["iters"], burn=st["burn"], b=st["b"], acceptance=m["acceptance"],
                             factorization_failures=m["factorization_failures"]))
    np.savez(os.path.join(HERE, "mcmc_n5000_heads.npz"), **chains)
    g["mcmc_n5000"] = dict(head=HEAD, theta_true=[1.0, 0.3, 0.05], seed=20261020, n_pred=300, chains=meta)
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
    import tlr_oracle as O
    S, y, Sp, yp = O.simulate_reference(5000)
    np.save(os.path.join(HERE, "sim_n5000_obs.npy"), np.column_stack([S, y]))
    np.save(os.path.join(HERE, "sim_n5000_holdout.npy"), np.column_stack([Sp, yp]))
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
