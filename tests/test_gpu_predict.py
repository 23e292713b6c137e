"""GPU kriging through the TLR factor (tlr_predict, tlr_forward/tlr_backward of codes/17) against the
oracle elementwise and against the reference's committed skill_true_theta (RMSE, CRPS, 90% coverage).
Tolerances: mean and sd 1e-9 relative to the oracle; skill scores 1e-8 absolute (the reference's
checker uses 1e-6 relative for fixed-theta quantities)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _field(n):
    a = np.load(os.path.join(GOLDEN, f"sim_n{n}_obs.npy"))
    return a[:, :2].copy(), a[:, 2].copy()


@pytest.fixture(scope="module")
def tlrgp():
    import tlrgp as m
    return m


@pytest.fixture(scope="module")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        return json.load(f)


def test_predict_vs_oracle_small(tlrgp):
    import tlr_oracle as O
    S, y = _field(2000)
    Sp = np.random.default_rng(5).random((77, 2))
    th = np.array([1.3, 0.2, 0.03])
    bnd = tlrgp.tile_bounds(2000, 250)
    print("case predict N=2000 b=250 adaptive:0.9", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    mu, sd = tlrgp.tlr_predict(S, y, Sp, th, bnd, fac)
    ll_o, fac_o, _ = O.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    mu_o, sd_o = O.tlr_predict(S, y, Sp, th, bnd, fac_o)
    np.testing.assert_allclose(mu, mu_o, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(sd, sd_o, rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("n", [10000, 40000])
def test_predict_skill_goldens(tlrgp, goldens, n):
    S, y = _field(n)
    hold = np.load(os.path.join(GOLDEN, f"sim_n{n}_holdout.npy"))
    th = np.array([1.0, 0.3, 0.05])
    bnd = tlrgp.tile_bounds(n, 250)
    for r in goldens["tlr_cholesky"][str(n)]["rules"]:
        if not r.get("completed") or "skill_true_theta" not in r:
            continue
        print(f"case predict N={n} b=250 {r['rule']}:{r['value']}", flush=True)
        ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, r["rule"], r["value"])
        mu, sd = tlrgp.tlr_predict(S, y, hold[:, :2], th, bnd, fac)
        sk = tlrgp.skill(hold[:, 2], mu, sd)
        for key in ("rmse", "crps", "cover90"):
            assert abs(sk[key] - r["skill_true_theta"][key]) < 1e-8, (r["rule"], r["value"], key, sk[key])


def test_predict_reuses_resident_factor(tlrgp):
    """tlr_predict with the factor tlr_loglik returned runs no second factorization (the factor handle), and
    gives the same values as a prediction that has to factorize again (another theta evaluated in between)."""
    S, y = _field(2000)
    Sp = np.random.default_rng(6).random((40, 2))
    th = np.array([1.1, 0.25, 0.04])
    bnd = tlrgp.tile_bounds(2000, 250)
    print("case predict factor handle N=2000 b=250", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    ctx = fac.ctx
    n0 = ctx.factorizations()
    mu, sd = tlrgp.tlr_predict(S, y, Sp, th, bnd, fac)
    assert ctx.factorizations() == n0
    ctx.loglik(np.array([0.9, 0.3, 0.05]), "adaptive", 0.9)   # slot 0 now holds another theta
    n1 = ctx.factorizations()
    mu2, sd2 = tlrgp.tlr_predict(S, y, Sp, th, bnd, fac)
    assert ctx.factorizations() == n1 + 1
    np.testing.assert_array_equal(mu, mu2)
    np.testing.assert_array_equal(sd, sd2)
    # solves reuse it too
    n2 = ctx.factorizations()
    z, o3 = ctx.solve(th, "adaptive", 0.9, y, "forward")
    assert ctx.factorizations() == n2
    assert abs(float(z @ z) - o3[2]) <= 1e-9 * o3[2]


def test_factor_outlives_context_cache_eviction():
    # tlr_loglik caches four contexts; a factor handle keeps its context alive after eviction, and a closed
    # context refuses calls with an error instead of passing a null handle to the library
    import tlrgp
    a = np.load(os.path.join(GOLDEN, "sim_n2000_obs.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    bnd = tlrgp.tile_bounds(2000, 250)
    th = np.array([1.0, 0.3, 0.05])
    Sp = S[:50] + 1e-3
    print("case factor handle across context-cache eviction N=2000 b=250", flush=True)
    ll0, fac0, _ = tlrgp.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    mu0, sd0 = tlrgp.tlr_predict(S, y, Sp, th, bnd, fac0)
    for k in range(1, 6):   # five other fields (shifted observations): the first context leaves the cache
        tlrgp.tlr_loglik(S, y + 0.01 * k, th, bnd, "adaptive", 0.9)
    mu1, sd1 = tlrgp.tlr_predict(S, y, Sp, th, bnd, fac0)
    np.testing.assert_array_equal(mu0, mu1)
    np.testing.assert_array_equal(sd0, sd1)
    fac0.ctx.close()
    with pytest.raises(tlrgp.TLRError, match="closed"):
        tlrgp.tlr_predict(S, y, Sp, th, bnd, fac0)
