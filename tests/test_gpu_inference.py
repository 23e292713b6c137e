"""Callers of the GPU likelihood against the reference's committed results:

* Nelder-Mead MLE at N=10k (codes/17_tlr_cholesky_pilot.py:380-404): the same simplex path, so the same
  theta_hat, objective, evaluation and failure counts, and the held-out skill at theta_hat;
* random-walk MH (codes/38_tlr_cholesky_mcmc.py:89-126) at N=5000: the first 1200 draws of the reference's
  chains (Robbins-Monro phase and the proposal re-shaping at iteration 1000 included), four adaptive
  chains and two fixed-rule chains advanced in lockstep with one batched evaluation per iteration;
* the gradient-based callers (L-BFGS-B, batched HMC): L-BFGS stationary on the Nelder-Mead optimum's
  ridge, HMC running over a batch of chains with bitwise checkpoint/resume on the device.
Tolerances: theta_hat 1e-7 relative, objective 1e-6 absolute, skill 1e-8, chain draws 1e-10 absolute."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        return json.load(f)


def _field(name):
    a = np.load(os.path.join(GOLDEN, name))
    return a[:, :2].copy(), a[:, 2].copy()


@pytest.mark.parametrize("rule,value", [("adaptive", 0.9), ("fixed", 1e-4)])
def test_mle_nelder_mead_matches_reference(goldens, rule, value):
    import tlrgp
    from tlrgp import inference as I
    g = goldens["mle"]["10000"]
    fit = next(f for f in g["fits"] if f["rule"] == rule and f["value"] == value)
    S, y = _field("sim_n10000_obs.npy")
    hold = np.load(os.path.join(GOLDEN, "sim_n10000_holdout.npy"))
    bnd = tlrgp.tile_bounds(10000, 250)
    x0 = np.log(np.array(g["theta_true"]) * np.array(g["x0_mult"]))
    print(f"case mle N=10000 b=250 {rule}:{value}", flush=True)
    r = I.fit_mle(S, y, bnd, rule, value, x0)
    print(r, flush=True)
    assert r["evals"] == fit["evals"] and r["failures"] == fit["failures"]
    np.testing.assert_allclose(r["theta_hat"], fit["theta_hat"], rtol=1e-7)
    assert abs(r["neg_loglik"] - fit["neg_loglik"]) < 1e-6
    ll, fac, _ = tlrgp.tlr_loglik(S, y, np.array(r["theta_hat"]), bnd, rule, value)
    mu, sd = tlrgp.tlr_predict(S, y, hold[:, :2], np.array(r["theta_hat"]), bnd, fac)
    sk = tlrgp.skill(hold[:, 2], mu, sd)
    for key in ("rmse", "crps", "cover90"):
        assert abs(sk[key] - fit["skill"][key]) < 1e-8, key


@pytest.mark.parametrize("rd,keys", [("adaptive_0.9", ("c1", "c2", "c3", "c4")), ("fixed_1e-4", ("c1", "c4"))])
def test_mh_chains_match_reference(goldens, rd, keys):
    import tlrgp
    from tlrgp import inference as I
    mc = goldens["mcmc_n5000"]
    heads = np.load(os.path.join(GOLDEN, "mcmc_n5000_heads.npz"))
    metas = [next(c for c in mc["chains"] if c["key"] == f"{rd}_{k}") for k in keys]
    S, y = _field("sim_n5000_obs.npy")
    m0 = metas[0]
    ctx = tlrgp.TLRContext(S, y, m0["b"], max_batch=len(keys))
    starts = np.log(np.array(mc["theta_true"]) * np.array([m["start_mult"] for m in metas]))
    print(f"case mh N=5000 b={m0['b']} {m0['rule']}:{m0['value']} chains {keys}", flush=True)
    res = I.run_chains_batched(I.batched_lik(ctx, m0["rule"], m0["value"]), starts,
                               [m["chain_seed"] for m in metas], mc["head"], m0["burn"])
    for m, r in zip(metas, res):
        ref = heads[m["key"]]
        np.testing.assert_allclose(r[0], ref, rtol=0, atol=1e-10, err_msg=m["key"])
    ctx.close()


def test_lbfgs_stationary_on_the_mle_ridge(goldens):
    import tlrgp
    from tlrgp import inference as I
    g = goldens["mle"]["10000"]
    fit = next(f for f in g["fits"] if f["rule"] == "adaptive")
    S, y = _field("sim_n10000_obs.npy")
    bnd = tlrgp.tile_bounds(10000, 250)
    x0 = np.log(np.array(g["theta_true"]) * np.array(g["x0_mult"]))
    print("case lbfgs N=10000 b=250 adaptive:0.9", flush=True)
    r = I.fit_mle_grad(S, y, bnd, "adaptive", 0.9, x0)
    print(r, flush=True)
    assert r["failures"] == 0
    # The log-likelihood is nearly flat along the ridge sigma2 / range^3 = const (the microergodic
    # parameter of Matern 3/2).  The frozen-rank gradient is the derivative of log L at fixed ranks; the
    # objective re-chooses ranks at every theta under a budget proportional to sigma2 * eta, and those
    # rank changes add a drift that the frozen-rank gradient does not see.  L-BFGS therefore ends its line
    # search on the ridge short of the derivative-free optimum (0.24 in log L on this field).
    th, tn = np.array(r["theta_hat"]), np.array(fit["theta_hat"])
    assert abs(th[0] / th[1] ** 3 / (tn[0] / tn[1] ** 3) - 1) < 0.03
    assert r["neg_loglik"] <= fit["neg_loglik"] + 0.5


def test_hmc_batched_runs_and_resumes(tmp_path):
    import tlrgp
    from tlrgp import inference as I
    S, y = _field("sim_n2000_obs.npy")
    M = 8
    ctx = tlrgp.TLRContext(S, y, 250, max_batch=M)
    lg = I.batched_loglik_grad(ctx, "adaptive", 0.9)
    starts = np.log([1.0, 0.3, 0.05]) + np.random.default_rng(1).normal(0, 0.05, (M, 3))
    print("case hmc N=2000 b=250 adaptive:0.9 chains 8", flush=True)
    r = I.hmc(lg, starts, 12, step=0.02, n_leapfrog=4, seed=5)
    assert np.all(np.isfinite(r["lls"]))
    assert np.mean(r["acceptance"]) > 0.3
    ck = str(tmp_path / "hmc.npz")
    n = dict(c=0)

    def interrupted(x):
        n["c"] += 1
        if n["c"] > 4 * 6 + 1:
            raise KeyboardInterrupt
        return lg(x)

    with pytest.raises(KeyboardInterrupt):
        I.hmc(interrupted, starts, 12, step=0.02, n_leapfrog=4, seed=5, ckpt=ck, ckpt_every=3)
    r2 = I.hmc(lg, starts, 12, step=0.02, n_leapfrog=4, seed=5, ckpt=ck, ckpt_every=3)
    assert np.array_equal(r["chain"], r2["chain"]) and np.array_equal(r["lls"], r2["lls"])
    # acceptance over all 12 iterations; latencies of the 6 saved iterations carried over, plus the resumed
    # run's own (its start is evaluated again)
    assert r2["acceptance"] == r["acceptance"]
    assert r2["evaluations"] == (1 + 6 * 4 + 1 + 6 * 4) * M
    ctx.close()
