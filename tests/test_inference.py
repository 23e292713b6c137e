"""Callers of the likelihood on the CPU: the MH sampler's steps against the reference's own chains (the
oracle as the likelihood), checkpoint/resume identity, the batched lockstep sampler against single chains,
HMC on a Gaussian target, and the chain diagnostics against closed forms."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

from tlrgp import inference as I


@pytest.fixture(scope="module")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        return json.load(f)


def _chain_meta(goldens, key):
    return next(c for c in goldens["mcmc_n5000"]["chains"] if c["key"] == key)


class OracleLik:
    """The reference's Lik (codes/38_tlr_cholesky_mcmc.py:39-56) over the oracle's tlr_loglik."""

    def __init__(self, S, y, b, rule, value):
        import tlr_oracle as O
        self.O, self.S, self.y, self.bnd, self.rule, self.value = O, S, y, O.tile_bounds(len(y), b), rule, value
        self.last = {}

    def __call__(self, theta):
        ll, _, _ = self.O.tlr_loglik(self.S, self.y, theta, self.bnd, self.rule, self.value)
        self.last = dict(ok=ll is not None)
        return -np.inf if ll is None else ll


def test_run_chain_reproduces_reference_chain_start(goldens):
    """Three MH iterations of chain 4 (adaptive 0.9, seed 104) with the oracle likelihood on the
    regenerated N=5000 field: the same proposals and accept decisions as the reference's chain
    (rows 1 and 2 are accepted moves), which pins the generator recipe, run_chain's use of the random
    stream and the likelihood along the way."""
    m = _chain_meta(goldens, "adaptive_0.9_c4")
    ref = np.load(os.path.join(GOLDEN, "mcmc_n5000_heads.npz"))["adaptive_0.9_c4"]
    a = np.load(os.path.join(GOLDEN, "sim_n5000_obs.npy"))
    lik = OracleLik(a[:, :2], a[:, 2], m["b"], m["rule"], m["value"])
    start = np.log(np.array(goldens["mcmc_n5000"]["theta_true"]) * np.array(m["start_mult"]))
    np.testing.assert_allclose(np.exp(start), m["start_theta"], rtol=1e-15)
    chain, lls, acc, fail, *_ = I.run_chain(lik, start, 3, m["burn"], m["chain_seed"], log=lambda s: None)
    np.testing.assert_allclose(chain, ref[:3], rtol=0, atol=1e-13)
    assert not np.array_equal(ref[1], ref[0]) and not np.array_equal(ref[2], ref[1])
    assert fail == 0


class FakeLik:
    """Gaussian log-density in log theta with a failure region (range > 0.36 fails like a non-PD
    factorization), deterministic."""
    mu = np.log([1.0, 0.3, 0.05])
    sd = np.array([0.3, 0.15, 0.25])

    def __init__(self, raise_after=None):
        self.calls = 0
        self.raise_after = raise_after
        self.last = {}

    def value(self, theta):
        lt = np.log(theta)
        ok = theta[1] <= 0.36
        return (-0.5 * np.sum(((lt - self.mu) / self.sd) ** 2) if ok else -np.inf), ok

    def __call__(self, theta):
        self.calls += 1
        if self.raise_after is not None and self.calls > self.raise_after:
            raise KeyboardInterrupt("simulated interruption")
        v, ok = self.value(np.asarray(theta))
        self.last = dict(ok=ok)
        return v

    def batch(self, thetas):
        out = [self.value(t) for t in np.atleast_2d(thetas)]
        return np.array([o[0] for o in out]), np.array([o[1] for o in out])


def test_run_chain_checkpoint_resume_is_bitwise(tmp_path):
    start = np.log([1.3, 0.24, 0.075])
    full = I.run_chain(FakeLik(), start, 700, 400, 7, log=lambda s: None)
    ck = str(tmp_path / "ck.npz")
    with pytest.raises(KeyboardInterrupt):
        I.run_chain(FakeLik(raise_after=350), start, 700, 400, 7, ckpt=ck, log=lambda s: None, ckpt_every=100)
    assert int(np.load(ck)["it"]) == 300
    res = I.run_chain(FakeLik(), start, 700, 400, 7, ckpt=ck, log=lambda s: None, ckpt_every=100)
    for a, b in zip(full[:6], res[:6]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_batched_chains_equal_single_chains(tmp_path):
    starts = np.log([[1.3, 0.24, 0.075], [0.6, 0.34, 0.025], [2.0, 0.15, 0.1], [0.8, 0.33, 0.006]])
    seeds = [101, 102, 103, 104]
    fl = FakeLik()
    batched = I.run_chains_batched(fl.batch, starts, seeds, 600, 300)
    for c in range(4):
        single = I.run_chain(FakeLik(), starts[c], 600, 300, seeds[c], log=lambda s: None)
        for a, b in zip(single[:6], batched[c][:6]):
            assert np.array_equal(np.asarray(a), np.asarray(b)), c
    assert sum(r[3] for r in batched) > 0   # the failure branch was taken
    ck = str(tmp_path / "ckb.npz")
    calls = dict(n=0)

    def interrupted(th):
        calls["n"] += 1
        if calls["n"] > 250:
            raise KeyboardInterrupt
        return fl.batch(th)

    with pytest.raises(KeyboardInterrupt):
        I.run_chains_batched(interrupted, starts, seeds, 600, 300, ckpt=ck, ckpt_every=100)
    resumed = I.run_chains_batched(fl.batch, starts, seeds, 600, 300, ckpt=ck, ckpt_every=100)
    for c in range(4):
        for k in range(6):
            assert np.array_equal(np.asarray(batched[c][k]), np.asarray(resumed[c][k]))


def test_hmc_gaussian_target_and_resume(tmp_path):
    mu = np.array([0.2, -1.2, -3.0])
    sd = np.array([0.3, 0.15, 0.25])

    def lg(x):
        return -0.5 * (((x - mu) / sd) ** 2).sum(1), -(x - mu) / sd ** 2

    starts = mu + np.random.default_rng(3).normal(0, 0.1, (16, 3))
    r = I.hmc(lg, starts, 400, step=0.05, n_leapfrog=8, seed=11)
    post = r["chain"][100:].reshape(-1, 3)
    assert np.all(np.abs(post.mean(0) - mu) < 0.1 * sd)
    np.testing.assert_allclose(post.std(0), sd, rtol=0.1)
    assert min(r["acceptance"]) > 0.5
    ck = str(tmp_path / "hmc.npz")
    n = dict(c=0)

    def lg_int(x):
        n["c"] += 1
        if n["c"] > 8 * 150 + 1:
            raise KeyboardInterrupt
        return lg(x)

    with pytest.raises(KeyboardInterrupt):
        I.hmc(lg_int, starts, 400, step=0.05, n_leapfrog=8, seed=11, ckpt=ck, ckpt_every=50)
    r2 = I.hmc(lg, starts, 400, step=0.05, n_leapfrog=8, seed=11, ckpt=ck, ckpt_every=50)
    assert np.array_equal(r["chain"], r2["chain"]) and np.array_equal(r["lls"], r2["lls"])
    # acceptance over all 400 iterations; the latencies of the 150 checkpointed iterations are carried over
    assert r2["acceptance"] == r["acceptance"]
    assert r2["evaluations"] == (1 + 150 * 8 + 1 + 250 * 8) * 16


def test_hmc_rejects_failed_evaluations():
    # an evaluator that has no result in part of the space (NaN log L and gradient, as for a theta beyond the
    # rank capacity or not positive definite): proposals landing there are rejected and the chains never
    # hold such a point
    mu = np.array([0.2, -1.2, -3.0])

    def lg(x):
        ll = -0.5 * (((x - mu) / 0.3) ** 2).sum(1)
        g = -(x - mu) / 0.09
        bad = x[:, 0] > 0.35
        return np.where(bad, np.nan, ll), np.where(bad[:, None], np.nan, g)

    starts = mu + np.random.default_rng(4).normal(0, 0.05, (8, 3))
    r = I.hmc(lg, starts, 200, step=0.1, n_leapfrog=5, seed=2)
    assert np.all(r["chain"][:, :, 0] <= 0.35) and np.all(np.isfinite(r["lls"]))
    assert 0.1 < np.mean(r["acceptance"]) < 1.0


def _ar1(phi, n, seed, m=1):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal((m, n))
    x = np.zeros((m, n))
    for t in range(1, n):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    return x


def test_ess_ar1_closed_form():
    """AR(1) with coefficient phi: integrated autocorrelation time (1 + phi) / (1 - phi)."""
    for phi in (0.0, 0.5, 0.8):
        x = _ar1(phi, 200000, 1)[0]
        want = 200000 * (1 - phi) / (1 + phi)
        assert abs(I.ess(x) / want - 1) < 0.06, phi
        assert abs(I.ess1(x) / want - 1) < 0.06, phi


def test_split_rhat_ess():
    x = _ar1(0.5, 4000, 2, m=4)
    rh, e = I.split_rhat_ess(x)
    assert abs(rh - 1) < 0.01
    assert abs(e / (4 * 4000 / 3) - 1) < 0.12
    x[0] += 2.0   # one chain off: R-hat flags it
    rh2, _ = I.split_rhat_ess(x)
    assert rh2 > 1.2
    s = I.pool_chains(0.1 * _ar1(0.3, 3000, 4, m=9).reshape(3, 3, 3000).transpose(0, 2, 1), burn=500)
    assert s["n_chains"] == 3 and len(s["split_rhat"]) == 3 and len(s["ci95"]) == 3


def test_concurrent_loglik_grad_splits_and_joins():
    """The multi-field evaluator hands each field its own rows, in order, and concatenates the results."""
    from tlrgp.inference import concurrent_loglik_grad
    seen = []

    def make(tag):
        def fn(x):
            seen.append((tag, x.shape[0]))
            return x.sum(1) + tag, x * tag
        return fn

    f = concurrent_loglik_grad([(make(1.0), 3), (make(10.0), 2)])
    x = np.arange(15, dtype=float).reshape(5, 3)
    ll, g = f(x)
    np.testing.assert_array_equal(ll, np.r_[x[:3].sum(1) + 1.0, x[3:].sum(1) + 10.0])
    np.testing.assert_array_equal(g, np.r_[x[:3] * 1.0, x[3:] * 10.0])
    assert sorted(seen) == [(1.0, 3), (10.0, 2)]
