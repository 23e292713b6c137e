"""GPU gradient of the TLR log-likelihood in (log sigma2, log range, log eta) at frozen ranks.

Checks: (1) against central finite differences of the CPU oracle's log L (h = 1e-5, ranks re-chosen,
which equals the frozen-rank derivative when no rank switches inside +-h; tolerance 1e-5 relative);
(2) against the dense FP64 analytic gradient with a tight fixed rule (eps = 1e-7, ranks up to 57 of the 64
columns a tile holds), where the TLR factor is close to the dense one (tolerance 1e-4 relative, the
target bound; the oracle's own distance is 3e-5); (3) with the adaptive rule the
distance to the dense gradient is the compression's own effect (reported, bounded by 5e-2); (4) the torch
autograd bridge returns the same gradient."""
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


def _oracle_fd(O, S, y, th, bnd, rule, value, h=1e-5):
    g = np.empty(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        lp = O.tlr_loglik(S, y, th * np.exp(e), bnd, rule, value)[0]
        lm = O.tlr_loglik(S, y, th * np.exp(-e), bnd, rule, value)[0]
        g[k] = (lp - lm) / (2 * h)
    return g


@pytest.mark.parametrize("theta", [(1.0, 0.3, 0.05), (1.5, 0.2, 0.02)])
def test_gradient_vs_oracle_fd(tlrgp, theta):
    import tlr_oracle as O
    S, y = _field(2000)
    th = np.array(theta)
    bnd = tlrgp.tile_bounds(2000, 250)
    print(f"case gradient N=2000 b=250 adaptive:0.9 theta={theta}", flush=True)
    ll, g, info = tlrgp.tlr_loglik_grad(S, y, th, bnd, "adaptive", 0.9)
    g_o = _oracle_fd(O, S, y, th, bnd, "adaptive", 0.9)
    print("gpu", g, "oracle fd", g_o, flush=True)
    np.testing.assert_allclose(g, g_o, rtol=1e-5, atol=1e-5)
    # the sigma2 component is the closed form 0.5 quad - N/2
    assert abs(g[0] - (0.5 * info["quad"] - 0.5 * 2000)) < 1e-9 * max(1.0, abs(g[0]))


def test_gradient_vs_dense_tight_rule(tlrgp):
    import tlr_oracle as O
    S, y = _field(2000)
    th = np.array([1.2, 0.25, 0.04])
    print("case gradient N=2000 fixed:1e-7 vs dense", flush=True)
    ll, g, info = tlrgp.tlr_loglik_grad(S, y, th, tlrgp.tile_bounds(2000, 250), "fixed", 1e-7)
    ll_d, g_d = O.loglik_dense_grad(S, y, th)
    print("gpu", ll, g, "dense", ll_d, g_d, flush=True)
    assert abs(ll - ll_d) < 2e-3
    np.testing.assert_allclose(g, g_d, rtol=1e-4, atol=1e-4)


def test_gradient_adaptive_distance_to_dense(tlrgp):
    import tlr_oracle as O
    S, y = _field(2000)
    th = np.array([1.0, 0.3, 0.05])
    ll, g, info = tlrgp.tlr_loglik_grad(S, y, th, tlrgp.tile_bounds(2000, 250), "adaptive", 0.9)
    ll_d, g_d = O.loglik_dense_grad(S, y, th)
    rel = np.abs(g - g_d) / np.maximum(1.0, np.abs(g_d))
    print("adaptive tau=0.9 gradient vs dense: relative", rel, flush=True)
    assert np.all(rel < 5e-2)


def test_autograd_bridge(tlrgp):
    torch = pytest.importorskip("torch")
    from tlrgp.autograd import tlr_loglik_torch
    S, y = _field(2000)
    ctx = tlrgp.TLRContext(S, y, 250, max_batch=2)
    lt = torch.tensor(np.log([[1.0, 0.3, 0.05], [1.5, 0.2, 0.02]]), dtype=torch.float64, requires_grad=True)
    ll = tlr_loglik_torch(lt, ctx)
    ll.sum().backward()
    ref = ctx.loglik_grad_batch(np.exp(lt.detach().numpy()), "adaptive", 0.9)
    np.testing.assert_allclose(ll.detach().numpy(), ref["loglik"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(lt.grad.numpy(), ref["grad"], rtol=1e-9, atol=1e-9)
