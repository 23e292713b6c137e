"""GPU parity of the TLR log-likelihood against the CPU oracle and the reference's golden values.

Tolerances: log L, log-det and quadratic form within 1e-6 * max(1, |v|) (the reference's own reproduction
tolerance for fixed-theta quantities, codes/29_check_reproduction.py:54); per-tile ranks equal to the
oracle's except where the oracle's tail lies within 1e-9 of the budget (counted and reported)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def _close(a, b, rtol=RTOL):
    return abs(a - b) <= rtol * max(1.0, abs(b))


def _golden_field(n):
    a = np.load(os.path.join(GOLDEN, f"sim_n{n}_obs.npy"))
    return a[:, :2].copy(), a[:, 2].copy()


@pytest.fixture(scope="module")
def tlrgp():
    import tlrgp as m
    print("device:", m.device_info(), flush=True)
    return m


@pytest.fixture(scope="module")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("rule,value", [("adaptive", 0.9), ("fixed", 1e-4), ("fixed", 1e-5)])
def test_small_field_vs_oracle(tlrgp, rule, value):
    import tlr_oracle as O
    S, y = _golden_field(2000)
    th = np.array([1.0, 0.3, 0.05])
    bnd = tlrgp.tile_bounds(len(y), 250)
    print(f"case N=2000 b=250 {rule}:{value}", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, rule, value)
    ll_o, _, info_o = O.tlr_loglik(S, y, th, bnd, rule, value)
    assert info["completed"] == info_o["completed"]
    np.testing.assert_array_equal(info["ranks_sigma"], info_o["ranks_sigma"])
    if info_o["completed"]:
        np.testing.assert_array_equal(info["ranks_factor"], info_o["ranks_factor"])
        assert _close(ll, ll_o), (ll, ll_o)
        assert _close(info["logdet"], info_o["logdet"])
        assert _close(info["quad"], info_o["quad"])
    else:
        assert info["failed_at"] == info_o["failed_at"]


def test_n10000_goldens(tlrgp, goldens):
    S, y = _golden_field(10000)
    g = goldens["tlr_cholesky"]["10000"]
    th = np.array([1.0, 0.3, 0.05])
    bnd = tlrgp.tile_bounds(len(y), 250)
    for r in g["rules"]:
        print(f"case N=10000 b=250 {r['rule']}:{r['value']}", flush=True)
        ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, r["rule"], r["value"])
        assert info["completed"] == r["completed"]
        assert info["stored_sigma"] == r["stored_sigma"]
        if r["completed"]:
            assert info["stored_factor"] == r["stored_factor"]
            assert info["max_rank_factor"] == r["max_rank_factor"]
            assert abs(info["mean_rank_factor"] - r["mean_rank_factor"]) < 1e-12
            assert _close(ll - g["dense_loglik"], r["loglik_err"], rtol=1e-5), (ll - g["dense_loglik"], r["loglik_err"])
        else:
            assert info["failed_at"] == r["failed_at_tile"]
    # separate TLR log-determinant golden of the same field (kl_tlr.json, adaptive tau=0.9)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    assert _close(info["logdet"], goldens["kl_tlr_n10000"]["logdet_approx"])


@pytest.mark.parametrize("theta", [(1.0, 0.3, 0.05), (0.5, 0.6, 0.01), (2.0, 0.1, 0.1), (1.3, 0.15, 0.02)])
def test_b500_vs_oracle(tlrgp, theta):
    """Tile size 500 (the benchmark tile), N=10,000 seed-0 field: ranks of every tile and log L against the
    oracle, over the corners of the benchmark's theta box (ranks up to ~50 exercise the wide sketch and the
    widest recompression class)."""
    import tlr_oracle as O
    S, y = _seed0_field(10000)
    th = np.array(theta)
    bnd = tlrgp.tile_bounds(len(y), 500)
    print(f"case N=10000 b=500 adaptive:0.9 theta={theta}", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    ll_o, _, info_o = O.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
    np.testing.assert_array_equal(info["ranks_sigma"], info_o["ranks_sigma"])
    assert info["completed"] == info_o["completed"]
    np.testing.assert_array_equal(info["ranks_factor"], info_o["ranks_factor"])
    assert _close(ll, ll_o), (ll, ll_o)


_SEED0 = {}


def _seed0_field(n):
    if n not in _SEED0:
        path = os.path.join(GOLDEN, f"seed0_n{n}.npy")
        if os.path.exists(path):
            a = np.load(path)
            _SEED0[n] = (a[:, :2].copy(), a[:, 2].copy())
        else:
            import tlr_oracle as O
            _SEED0[n] = O.simulate_seed0(n)
    return _SEED0[n]


def test_n40000_b250_golden(tlrgp, goldens):
    """Full-size reference run (N=40,000, b=250, adaptive 0.9): log L, stored factor and ranks against the
    committed results of codes/17 (tlr_cholesky_n40000.json)."""
    S, y = _golden_field(40000)
    g = goldens["tlr_cholesky"]["40000"]
    r = [x for x in g["rules"] if x["rule"] == "adaptive"][0]
    print("case N=40000 b=250 adaptive:0.9", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, np.array([1.0, 0.3, 0.05]), tlrgp.tile_bounds(40000, 250), "adaptive", 0.9)
    assert info["completed"]
    assert info["stored_sigma"] == r["stored_sigma"]
    assert info["stored_factor"] == r["stored_factor"]
    assert info["max_rank_factor"] == r["max_rank_factor"]
    assert _close(ll, r["loglik"]), (ll, r["loglik"])


def test_n40000_b500_vs_oracle(tlrgp):
    """The benchmark configuration (N=40,000, b=500, adaptive 0.9) at theta=(1, 0.3, 0.05) on the seed-0
    field: every initial and final tile rank and log L / log-det / quadratic form against the oracle's
    stored result (tests/golden/make_seed0_oracle.py)."""
    path = os.path.join(GOLDEN, "seed0_n40000_b500_oracle.npz")
    if not os.path.exists(path):
        pytest.skip("oracle fixture not generated")
    ref = np.load(path)
    S, y = _seed0_field(40000)
    print("case N=40000 b=500 adaptive:0.9 seed0", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, ref["theta"], tlrgp.tile_bounds(40000, 500), "adaptive", 0.9)
    np.testing.assert_array_equal(info["ranks_sigma"], ref["ranks_sigma"])
    np.testing.assert_array_equal(info["ranks_factor"], ref["ranks_factor"])
    assert _close(ll, float(ref["loglik"])), (ll, float(ref["loglik"]))
    assert _close(info["logdet"], float(ref["logdet"]))
    assert _close(info["quad"], float(ref["quad"]))


def test_n40000_b500_lhs_thetas_vs_oracle(tlrgp):
    """Three of the benchmark's own Latin-hypercube theta (N=40,000, b=500, adaptive 0.9, seed-0 field), in one
    batch: every initial and final tile rank identical to the oracle's and log L / log-det / quadratic form
    within 1e-6 relative (tests/golden/make_seed0_lhs_oracle.py)."""
    path = os.path.join(GOLDEN, "seed0_n40000_b500_lhs_oracle.npz")
    if not os.path.exists(path):
        pytest.skip("oracle fixture not generated")
    ref = np.load(path)
    S, y = _seed0_field(40000)
    th = ref["theta"]
    print(f"case N=40000 b=500 adaptive:0.9 seed0, LHS rows {ref['rows'].tolist()}", flush=True)
    ctx = tlrgp.TLRContext(S, y, 500, max_batch=len(th))
    r = ctx.loglik_batch(th, "adaptive", 0.9)
    ctx.close()
    assert r["completed"].all()
    for m in range(len(th)):
        np.testing.assert_array_equal(r["ranks_sigma"][m], ref["ranks_sigma"][m])
        np.testing.assert_array_equal(r["ranks_factor"][m], ref["ranks_factor"][m])
        assert _close(r["loglik"][m], float(ref["loglik"][m])), (m, r["loglik"][m], float(ref["loglik"][m]))
        assert _close(r["logdet"][m], float(ref["logdet"][m]))
        assert _close(r["quad"][m], float(ref["quad"][m]))


@pytest.mark.parametrize("n", [60000, 100000])
def test_largest_reference_runs(tlrgp, goldens, n):
    """The reference's two largest runs (b=250, theta=(1, 0.3, 0.05); N=100,000 is 79,800 off-diagonal tiles):
    the adaptive rule's log L, stored values and mean/max factor rank, and the fixed 1e-4 rule's failing tile,
    against tlr_cholesky_n{60000,100000}.json."""
    S, y = _golden_field(n)
    th = np.array([1.0, 0.3, 0.05])
    bnd = tlrgp.tile_bounds(n, 250)
    for r in goldens["tlr_cholesky"][str(n)]["rules"]:
        print(f"case N={n} b=250 {r['rule']}:{r['value']}", flush=True)
        ll, fac, info = tlrgp.tlr_loglik(S, y, th, bnd, r["rule"], r["value"])
        assert info["completed"] == r["completed"]
        assert info["stored_sigma"] == r["stored_sigma"]
        if r["completed"]:
            assert info["stored_factor"] == r["stored_factor"]
            assert info["max_rank_factor"] == r["max_rank_factor"]
            assert abs(info["mean_rank_factor"] - r["mean_rank_factor"]) < 1e-12
            assert _close(ll, r["loglik"]), (ll, r["loglik"])
        else:
            assert info["failed_at"] == r["failed_at_tile"]
        fac = None
    # release the large context (the facade caches up to four)
    for c in list(tlrgp.likelihood._CTX_CACHE.values()):
        c.close()
    tlrgp.likelihood._CTX_CACHE.clear()


@pytest.mark.parametrize("tag", ["nu15_r03_e005", "nu15_r03_e05", "nu15_r1_e005", "nu15_r3_e005"])
def test_screening_fields_at_tlr_mle(tlrgp, goldens, tag):
    """The reference's TLR maximum-likelihood points on the four Matern-3/2 fields of its screening study
    (codes/27_inference_headtohead.py: 17's log L with the adaptive rule at theta_hat): same N=10k sites,
    each field's own observations, ranges 0.3-3, nuggets 0.03-0.44."""
    g = goldens["screening_tlr_mle"][tag]
    S, _ = _golden_field(10000)
    y = np.load(os.path.join(GOLDEN, f"screening_{tag}_z.npy"))
    th = np.array(g["theta_hat"])
    print(f"case screening field {tag} N=10000 b={g['b']} adaptive:{g['tau']} theta_hat={g['theta_hat']}", flush=True)
    ll, fac, info = tlrgp.tlr_loglik(S, y, th, tlrgp.tile_bounds(10000, g["b"]), "adaptive", g["tau"])
    print("log L", ll, "reference", g["tlr_loglik"], flush=True)
    assert info["completed"]
    # the TLR error against the dense log L, as for the reference's own N=10k runs
    assert _close(ll - g["exact_loglik"], g["tlr_loglik"] - g["exact_loglik"], rtol=1e-5), (ll, g["tlr_loglik"])
