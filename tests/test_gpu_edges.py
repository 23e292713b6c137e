"""Edge shapes of the tile grid against the oracle: a short last tile, tile sizes that are not multiples of
the 32/64-row blocks of the kernels, a two-tile problem, tiles narrower than the 96-column second round (b = 60:
45 of 66 tiles are redone at k = b = 60 columns, ranks up to 49; b = 40: a single round at k = b), and both
rank rules; initial and final ranks identical, log L, log-det and quadratic form within 1e-6 * max(1, |v|)
(the reference checker's tolerance).  Also the rank capacity: a theta whose tile needs more than 64 columns
fails alone in its batch."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CASES = [  # (N, b, theta, rule, value)
    (2345, 250, (1.0, 0.3, 0.05), "adaptive", 0.9),
    (1000, 300, (0.7, 0.15, 0.02), "adaptive", 0.9),
    (777, 128, (1.5, 0.45, 0.08), "fixed", 1e-5),
    (300, 250, (1.0, 0.2, 0.1), "adaptive", 0.5),
    (1500, 500, (2.0, 0.6, 0.01), "fixed", 1e-4),
    (700, 60, (1.0, 0.25, 0.05), "fixed", 1e-10),
    (500, 40, (0.8, 0.2, 0.02), "fixed", 1e-6),
]


def _field(n, seed=11):
    import tlr_oracle as O
    rng = np.random.default_rng(seed)
    S = rng.random((n, 2))
    S = S[O.morton_order(S)]
    y = np.linalg.cholesky(O.cov_full(S, np.array([1.0, 0.3, 0.05]))) @ rng.standard_normal(n)
    return S, y


def _close(a, b):
    return abs(a - b) <= 1e-6 * max(1.0, abs(b))


@pytest.mark.parametrize("n,b,theta,rule,value", CASES)
def test_edge_shapes_vs_oracle(n, b, theta, rule, value):
    import tlrgp
    import tlr_oracle as O
    S, y = _field(n)
    bnd = tlrgp.tile_bounds(n, b)
    th = np.array(theta)
    print(f"case edge N={n} b={b} {rule}:{value} theta={theta}", flush=True)
    ll, _, info = tlrgp.tlr_loglik(S, y, th, bnd, rule, value)
    ll_o, _, info_o = O.tlr_loglik(S, y, th, bnd, rule, value)
    np.testing.assert_array_equal(info["ranks_sigma"], info_o["ranks_sigma"])
    assert info["completed"] == info_o["completed"]
    if info_o["completed"]:
        np.testing.assert_array_equal(info["ranks_factor"], info_o["ranks_factor"])
        assert _close(ll, ll_o), (ll, ll_o)
        assert _close(info["logdet"], info_o["logdet"]) and _close(info["quad"], info_o["quad"])
    else:
        assert info["failed_at"] == info_o["failed_at"]


def test_rank_capacity_fails_only_its_theta():
    # N=10k, b=500, range 0.05, nugget 0.05: a tile's initial rank is 66 (results/joint_sweep), beyond the
    # 64-column capacity.  In a batch that theta alone has no result; alone it raises.
    import tlrgp
    a = np.load(os.path.join(GOLDEN, "sim_n10000_obs.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    capped, fine = [1.0, 0.05, 0.05], [1.0, 0.3, 0.05]
    ctx = tlrgp.TLRContext(S, y, 500, max_batch=10)
    print("case rank capacity N=10000 b=500 adaptive:0.9", flush=True)
    r = ctx.loglik_batch(np.array([capped, fine]), "adaptive", 0.9)
    assert r["rank_capped"].tolist() == [True, False] and r["completed"].tolist() == [False, True]
    assert np.isnan(r["loglik"][0]) and r["failed_at"][0] == -1
    r1 = ctx.loglik_batch(np.array([fine]), "adaptive", 0.9)
    np.testing.assert_allclose(r1["loglik"][0], r["loglik"][1], rtol=1e-12)
    with pytest.raises(tlrgp.TLRError, match="capacity"):
        ctx.loglik_batch(np.array([capped]), "adaptive", 0.9)
    g = ctx.loglik_grad_batch(np.array([capped, fine]), "adaptive", 0.9)
    g1 = ctx.loglik_grad_batch(np.array([fine]), "adaptive", 0.9)
    assert g["rank_capped"].tolist() == [True, False] and np.all(np.isnan(g["grad"][0]))
    np.testing.assert_allclose(g["grad"][1], g1["grad"][0], rtol=1e-9)
    ctx.close()
    # the one-theta samplers treat it as a failed factorization
    bnd = tlrgp.tile_bounds(10000, 500)
    lik = tlrgp.Lik(S, y, bnd, "adaptive", 0.9)
    assert lik(np.array(capped)) == -np.inf and lik.last["rank_capped"]
    trace = dict(evals=0, failures=0)
    assert tlrgp.neg_loglik_tlr(np.log(capped), S, y, bnd, "adaptive", 0.9, trace) == 1e12
    assert trace == dict(evals=1, failures=1, rank_capped=1)
