"""The gradient configuration: N=20,000, b=250, adaptive tau=0.9, six fields (range in {0.1, 0.3, 1.0} x
eta in {0.01, 0.05}, sigma2 = 1, seed-0 sites), each at its own theta, against the oracle's frozen-rank
central differences (h = 1e-5) and the dense FP64 analytic gradient (tests/golden/make_grad_n20000.py);
the same gradient with the reduced-precision arithmetic; and the observed information (negative Hessian in
log theta by differencing gradients) against the reference's dense standard errors
(codes/15_pilot_standard_errors.py; results/pilot_standard_errors_n{10000,20000}.json)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "grad_n20000_b250.npz")


@pytest.fixture(scope="module")
def tl():
    import tlrgp
    return tlrgp


@pytest.fixture(scope="module")
def fx():
    if not os.path.exists(FIX):
        pytest.skip("gradient fixture not generated")
    return np.load(FIX)


def test_gradient_six_fields(tl, fx):
    S = fx["S"]
    worst_dense = 0.0
    for f, (r, e) in enumerate(fx["fields"]):
        th = np.array([1.0, r, e])
        y = fx["ys"][f]
        print(f"case gradient N=20000 b=250 adaptive:0.9 field range={r} eta={e}", flush=True)
        ctx = tl.TLRContext(S, y, 250, max_batch=1)
        res = ctx.loglik_grad_batch(th[None, :], "adaptive", 0.9, h=1e-5, ranks=False)
        assert bool(res["completed"][0]) == bool(fx["completed"][f])
        if not fx["completed"][f]:
            ctx.close()
            continue
        g, g_fd, g_d = res["grad"][0], fx["g_fd"][f], fx["g_dense"][f]
        ll = res["loglik"][0]
        assert abs(ll - fx["ll_tlr"][f]) <= 1e-6 * max(1.0, abs(fx["ll_tlr"][f])), (ll, fx["ll_tlr"][f])
        # frozen-rank central differences of the oracle: the same derivative
        np.testing.assert_allclose(g, g_fd, rtol=1e-5, atol=1e-5 * np.abs(g_fd).max())
        # relative to the gradient's largest component: a component near zero (eta at the true theta) is a
        # small difference of large terms, and the rule's approximation moves it by a share of those terms
        rel = np.abs(g - g_d) / max(1.0, float(np.abs(g_d).max()))
        worst_dense = max(worst_dense, float(rel.max()))
        # reduced-precision arithmetic (frozen ranks and arithmetic) against FP64 with the same difference step
        # (1e-3: the arithmetic's error, not the step's truncation error, is what is compared), and against the
        # FP64 gradient at the default step (both errors together).  An FP32 error delta in log L moves a
        # difference quotient by up to delta / h in every component (delta ~ 1e-5 here), so the distances are
        # taken relative to the gradient's largest component, as for the dense gradient above.
        r32 = ctx.loglik_grad_batch(th[None, :], "adaptive", 0.9, h=1e-3, precision="fp32", ranks=False)
        r64h = ctx.loglik_grad_batch(th[None, :], "adaptive", 0.9, h=1e-3, precision="fp64", ranks=False)
        scale = max(1.0, float(np.abs(g).max()))
        d32 = np.abs(r32["grad"][0] - r64h["grad"][0]) / scale
        dstep = np.abs(r32["grad"][0] - g) / scale
        print(f"  gpu {g} oracle fd {g_fd} dense {g_d}: vs dense {rel}; fp32 vs fp64 (h=1e-3) {d32}; "
              f"fp32 (h=1e-3) vs fp64 (h=1e-5) {dstep}; dlogL fp32 {r32['loglik'][0] - ll:.2e}", flush=True)
        assert abs(r32["loglik"][0] - ll) <= 1e-2
        assert np.all(d32 <= 1e-3)
        assert np.all(dstep <= 1e-2)
        ctx.close()
    # the adaptive rule's own approximation separates the TLR gradient from the dense one
    print("largest relative distance to the dense gradient:", worst_dense, flush=True)
    assert worst_dense < 5e-2


def _se_goldens():
    with open(os.path.join(GOLDEN, "standard_errors.json")) as f:
        return json.load(f)


def _field(n):
    a = np.load(os.path.join(GOLDEN, f"sim_n{n}_obs.npy"))
    return a[:, :2].copy(), a[:, 2].copy()


def test_observed_information_n10000(tl):
    """At the reference's dense MLE on its N=10k field (b=250, adaptive 0.9): the observed information by
    differencing GPU gradients against the reference's dense one (second differences of dense log L, h=0.02).
    The TLR approximation moves the curvature (the reference's own assembled-adaptive fit differs from its
    dense one by up to 16 % on the diagonal), so the diagonal is held to 25 % and the off-diagonal signs must
    agree; the reference's own scheme (second differences of log L) on the GPU agrees with the gradient
    scheme to 10 %."""
    from tlrgp.inference import observed_information
    g = _se_goldens()["10000"]
    S, y = _field(10000)
    th = np.array(g["dense"]["theta_hat"])
    bnd = tl.tile_bounds(10000, 250)
    res = observed_information(S, y, th, bnd, "adaptive", 0.9, h=g["h"], method="gradient")
    res2 = observed_information(S, y, th, bnd, "adaptive", 0.9, h=g["h"], method="loglik")
    ref = np.array(g["dense"]["info_log"])
    print("info (gradient differences)", res["info_log"].tolist(), "\ninfo (log L differences)",
          res2["info_log"].tolist(), "\nreference dense", ref.tolist(), "\nse_log", res["se_log"],
          "reference", g["dense"]["se_log"], flush=True)
    d = np.diag(res["info_log"])
    np.testing.assert_allclose(d, np.diag(ref), rtol=0.25)
    np.testing.assert_array_equal(np.sign(res["info_log"]), np.sign(ref))
    np.testing.assert_allclose(np.diag(res2["info_log"]), d, rtol=0.1)


def test_standard_errors_n20000(tl):
    """At the reference's rounded dense MLE on its committed N=20k field: the observed information by
    differencing GPU gradients against the dense FP64 one on the same field (tests/golden/make_info_n20000.py,
    the reference's scheme), held as at N=10k (diagonal 25 %, signs), and the gradient against the dense
    analytic one on the scale N/2 of its terms.  The reference's own N=20k standard errors were computed on
    data regenerated from the seed and are printed for comparison only."""
    from tlrgp.inference import observed_information
    path = os.path.join(GOLDEN, "info_n20000_dense.npz")
    if not os.path.exists(path):
        pytest.skip("dense information fixture not generated")
    dense = np.load(path)
    g = _se_goldens()["20000"]
    S, y = _field(20000)
    th = dense["theta"]
    bnd = tl.tile_bounds(20000, 250)
    ll, grad, info = tl.tlr_loglik_grad(S, y, th, bnd, "adaptive", 0.9)
    res = observed_information(S, y, th, bnd, "adaptive", 0.9, h=float(dense["h"]), method="gradient")
    print("gradient", grad, "dense", dense["gradient"], "\ninfo", res["info_log"].tolist(), "\ndense",
          dense["info_log"].tolist(), "\nse_log", res["se_log"], "dense", dense["se_log"], "reference (regenerated "
          "data)", g["dense"]["se_log"], flush=True)
    assert np.all(np.abs(grad - dense["gradient"]) <= 1e-3 * 20000 / 2)
    np.testing.assert_allclose(np.diag(res["info_log"]), np.diag(dense["info_log"]), rtol=0.25)
    np.testing.assert_array_equal(np.sign(res["info_log"]), np.sign(dense["info_log"]))
