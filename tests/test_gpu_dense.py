"""The dense FP64 log-likelihood on the GPU (tlr_dense_loglik: loglik_dense of codes/11_tlr_rank_rules.py:112-120,
the exact likelihood the TLR one approximates): against the oracle at N=2000 (log L, log-determinant, quadratic
form at several theta), and against the reference's dense values at N=10k
(dense_loglik_true, kl_tlr.json#logdet_exact) and N=40k (dense_loglik_true)."""
import json
import os

import numpy as np
import pytest

import tlr_oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _field(n):
    a = np.load(os.path.join(GOLDEN, f"sim_n{n}_obs.npy"))
    return a[:, :2].copy(), a[:, 2].copy()


def test_dense_loglik_vs_oracle_n2000():
    import tlrgp
    S, y = _field(2000)
    th = np.array([[1.0, 0.3, 0.05], [0.6, 0.12, 0.2], [1.8, 0.5, 0.01]])
    ctx = tlrgp.TLRContext(S, y, 250, max_batch=1)
    print("case dense log L N=2000, 3 theta", flush=True)
    r = ctx.dense_loglik_batch(th)
    assert r["completed"].all()
    for s in range(len(th)):
        ll_o, cho = O.loglik_dense(O.cov_full(S, th[s]), y)
        quad_o = float(y @ cho_solve_(cho, y))
        logdet_o = 2.0 * float(np.log(np.diag(cho[0])).sum())
        assert abs(r["loglik"][s] - ll_o) <= 1e-10 * abs(ll_o), (s, r["loglik"][s], ll_o)
        assert abs(r["logdet"][s] - logdet_o) <= 1e-10 * abs(logdet_o)
        assert abs(r["quad"][s] - quad_o) <= 1e-10 * abs(quad_o)
    ctx.close()


def cho_solve_(cho, y):
    from scipy.linalg import cho_solve
    return cho_solve(cho, y)


@pytest.mark.parametrize("n", [10000, 40000])
def test_dense_loglik_reference_values(n):
    import tlrgp
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        g = json.load(f)
    S, y = _field(n)
    ctx = tlrgp.TLRContext(S, y, 250 if n == 10000 else 500, max_batch=1)
    print(f"case dense log L N={n} at theta_true", flush=True)
    r = ctx.dense_loglik_batch(np.array([[1.0, 0.3, 0.05]]))
    want = g["tlr_cholesky"][str(n)]["dense_loglik"]
    assert abs(r["loglik"][0] - want) <= 1e-10 * abs(want), (r["loglik"][0], want)
    if n == 10000:
        ld = g["kl_tlr_n10000"]["logdet_exact"]
        assert abs(r["logdet"][0] - ld) <= 1e-10 * abs(ld), (r["logdet"][0], ld)
    ctx.close()
