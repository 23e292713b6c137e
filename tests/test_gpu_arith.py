"""Reduced-precision arithmetic of the factorization's off-diagonal work (TLR_PREC_FP32 and TLR_PREC_BF16:
products, recompression QRs and core SVDs in FP32, the BF16 mode with bfloat16 product operands) against the
plain FP64 path and the dense FP64 oracle.  No counterpart in the reference (its arithmetic is FP64); the
bound is the north star's |delta log L| <= 1e-2 against plain FP64, and the completion (PD) flags must agree.
"""
import os

import numpy as np
import pytest

import tlr_oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

# corners of the benchmark's box and a small-range, small-nugget theta (b = 250 only: at b = 500 it needs more
# than the 64-column rank capacity in every arithmetic)
THETAS = np.array([[1.0, 0.3, 0.05], [0.5, 0.6, 0.01], [2.0, 0.1, 0.1], [1.5, 0.12, 0.011]])
THETAS_B500 = np.array([[1.0, 0.3, 0.05], [0.5, 0.6, 0.01], [2.0, 0.1, 0.1], [0.7, 0.45, 0.02]])
DLOGL = 1e-2   # |log L(reduced) - log L(FP64)|


@pytest.fixture(scope="module")
def tl():
    import tlrgp
    return tlrgp


def _field(n):
    if n in (10000, 40000):
        a = np.load(os.path.join(GOLDEN, f"seed0_n{n}.npy"))
        return a[:, :2].copy(), a[:, 2].copy()
    return O.simulate_seed0(n, seed=0)


@pytest.mark.parametrize("n,b", [(2000, 250), (2345, 250), (10000, 500)])
def test_reduced_precision_vs_fp64(tl, n, b):
    S, y = _field(n)
    th = THETAS if b <= 256 else THETAS_B500
    ctx = tl.TLRContext(S, y, b, max_batch=len(th))
    r64 = ctx.loglik_batch(th, "adaptive", 0.9, "fp64")
    c64 = [ctx.arith_counts(s) for s in range(len(th))]
    assert all(c["fp32"] == 0 and c["bf16"] == 0 for c in c64)
    for prec in ("fp32", "bf16"):
        r = ctx.loglik_batch(th, "adaptive", 0.9, prec)
        counts = [ctx.arith_counts(s) for s in range(len(th))]
        print(f"N={n} b={b} {prec}: dlogL={r['loglik'] - r64['loglik']} counts={counts}", flush=True)
        np.testing.assert_array_equal(r["completed"], r64["completed"])
        ok = r64["completed"]
        assert np.all(np.abs(r["loglik"][ok] - r64["loglik"][ok]) <= DLOGL), (prec, r["loglik"] - r64["loglik"])
        # initial compression (FP64 in every mode): identical ranks
        np.testing.assert_array_equal(r["ranks_sigma"], r64["ranks_sigma"])
        # at the default guard some pairs still qualify for FP32 on these fields
        if prec == "fp32":
            assert sum(c["fp32"] for c in counts) > 0
        # every recompressed pair ran in exactly one arithmetic
        for c in counts:
            assert c["fp64"] + c["fp32"] + c["bf16"] > 0 or n <= b


def test_reduced_precision_vs_dense(tl):
    """Error of each mode against the dense FP64 log-likelihood at N=2000 (the TLR approximation's own error
    dominates; the reduced-precision modes must not move it by more than the bound)."""
    S, y = _field(2000)
    ctx = tl.TLRContext(S, y, 250, max_batch=1)
    th = THETAS[0]
    dense = O.loglik_dense_parts(S, y, th)["loglik"]
    err = {}
    for prec in ("fp64", "fp32", "bf16"):
        r = ctx.loglik_batch(th[None, :], "adaptive", 0.9, prec)
        err[prec] = float(r["loglik"][0] - dense)
    print("error against dense FP64:", err, flush=True)
    assert abs(err["fp64"]) < 1.0
    for prec in ("fp32", "bf16"):
        assert abs(err[prec] - err["fp64"]) <= DLOGL


def test_reduced_precision_fixed_rule(tl):
    """Fixed relative rule: eps = 1e-4 is 800 FP32 unit roundoffs, so with a guard of 64 most pairs run in
    FP32 (the default guard, 16384, asks for 2e-3)."""
    S, y = _field(2000)
    ctx = tl.TLRContext(S, y, 250, max_batch=1)
    ctx.set_tunable("guard", 64.0)
    th = THETAS[:1]
    r64 = ctx.loglik_batch(th, "fixed", 1e-4, "fp64")
    r32 = ctx.loglik_batch(th, "fixed", 1e-4, "fp32")
    c = ctx.arith_counts(0)
    assert c["fp32"] > c["fp64"]
    assert r32["completed"][0] == r64["completed"][0]
    if r64["completed"][0]:
        assert abs(r32["loglik"][0] - r64["loglik"][0]) <= DLOGL


def test_guard_tunable_moves_pairs(tl):
    """A smaller guard sends more pairs to FP32; every guard keeps each pair in exactly one arithmetic."""
    S, y = _field(2000)
    ctx = tl.TLRContext(S, y, 250, max_batch=1)
    th = THETAS[:1]
    fr = []
    for g in (64.0, 16384.0, 1e9):
        ctx.set_tunable("guard", g)
        assert ctx.get_tunable("guard") == g
        ctx.loglik_batch(th, "adaptive", 0.9, "fp32")
        c = ctx.arith_counts(0)
        fr.append(c["fp32"])
    assert fr[0] >= fr[1] >= fr[2]
    assert fr[0] > fr[2]


def test_reduced_precision_repeatable(tl):
    """Same inputs, same results bit for bit (fixed-order reductions; the guard's decisions are per pair)."""
    S, y = _field(2000)
    ctx = tl.TLRContext(S, y, 250, max_batch=4)
    a = ctx.loglik_batch(THETAS, "adaptive", 0.9, "fp32")
    b = ctx.loglik_batch(THETAS, "adaptive", 0.9, "fp32")
    np.testing.assert_array_equal(a["loglik"], b["loglik"])
    np.testing.assert_array_equal(a["ranks_factor"], b["ranks_factor"])


def test_mixed_storage_rule_needs_fp64(tl):
    S, y = _field(2000)
    ctx = tl.TLRContext(S, y, 250, max_batch=1)
    with pytest.raises(tl.TLRError):
        ctx.loglik_batch(THETAS[:1], "mixed", 0.9, "fp32")
