"""Multi-right-hand-side solves with the GPU TLR factor (tlr_forward 17:230-240, tlr_backward 17:243-256)
against the oracle, and the KL certificate of codes/31_kl_tlr.py against the reference's kl_tlr.json (the N=10k
field and the three other Matern-3/2 fields of its screening study).

Tolerances: solves 1e-9 relative to the oracle's (the factors agree to rounding); the KL's trace and
log-determinants 1e-9 relative, the KL itself (a difference 1e-5 the size of its terms) 1e-4 relative."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _field(n):
    a = np.load(os.path.join(GOLDEN, f"sim_n{n}_obs.npy"))
    return a[:, :2].copy(), a[:, 2].copy()


def test_solves_vs_oracle():
    import tlrgp
    import tlr_oracle as O
    S, y = _field(2000)
    th = np.array([0.8, 0.35, 0.06])
    bnd = tlrgp.tile_bounds(2000, 250)
    B = np.random.default_rng(9).standard_normal((2000, 5))
    print("case solve N=2000 b=250 adaptive:0.9", flush=True)
    Ld, Lo, _ = O.tlr_cholesky(S, th, bnd, "adaptive", 0.9)
    zf = O.tlr_forward(Ld, Lo, bnd, B)
    zb = O.tlr_backward(Ld, Lo, bnd, B)
    zs = O.tlr_backward(Ld, Lo, bnd, zf)
    np.testing.assert_allclose(tlrgp.tlr_forward(S, y, th, bnd, "adaptive", 0.9, B), zf, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(tlrgp.tlr_backward(S, y, th, bnd, "adaptive", 0.9, B), zb, rtol=1e-9, atol=1e-9)
    ctx = tlrgp.TLRContext(S, y, 250)
    X, o3 = ctx.solve(th, "adaptive", 0.9, B[:, 2], "both")
    np.testing.assert_allclose(X, zs[:, 2], rtol=1e-9, atol=1e-9)
    ctx.close()


def test_kl_tlr_matches_reference():
    import tlrgp
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        g = json.load(f)["kl_tlr_n10000"]
    S, _ = _field(10000)
    print("case kl N=10000 b=250 adaptive:0.9", flush=True)
    r = tlrgp.kl_tlr(S, np.array([1.0, 0.3, 0.05]), tlrgp.tile_bounds(10000, 250), "adaptive", g["tau"])
    print(r, flush=True)
    for key in ("trace", "logdet_exact", "logdet_approx"):
        assert abs(r[key] - g[key]) <= 1e-9 * abs(g[key]), key
    assert abs(r["kl"] - g["kl"]) <= 1e-4 * abs(g["kl"])


@pytest.mark.parametrize("tag", ["nu15_r03_e05", "nu15_r1_e005", "nu15_r3_e005"])
def test_kl_tlr_screening_fields(tag):
    # the other Matern-3/2 fields of the reference's screening study (same sites, larger nugget or range)
    import tlrgp
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        g = json.load(f)["kl_tlr_screening"][tag]
    S, _ = _field(10000)
    print(f"case kl N=10000 b={g['b']} adaptive:{g['tau']} field {tag} theta={g['theta']}", flush=True)
    r = tlrgp.kl_tlr(S, np.array(g["theta"]), tlrgp.tile_bounds(10000, g["b"]), "adaptive", g["tau"])
    print(r, flush=True)
    for key in ("trace", "logdet_exact", "logdet_approx"):
        assert abs(r[key] - g[key]) <= 1e-9 * abs(g[key]), (key, r[key], g[key])
    assert abs(r["kl"] - g["kl"]) <= 1e-4 * abs(g["kl"]), (r["kl"], g["kl"])
