"""Gradient fixtures at N=20,000, b=250, adaptive tau=0.9 (the gradient configuration): six fields, one per
(range, eta) in {0.1, 0.3, 1.0} x {0.01, 0.05} with sigma2 = 1, all on the seed-0 sites (S ~ U[0,1]^2 from
default_rng(0), Morton order; y drawn by the dense FP64 Cholesky of Sigma(1, range, eta), simulate_seed0).
At each field's own theta: the dense FP64 log L and analytic gradient (loglik_dense_grad) and the oracle's
TLR log L with its frozen-rank central differences in log range and log eta (h = 1e-5; the log sigma2
component is the closed form quad / 2 - N / 2), written to grad_n20000_b250.npz with the sites and the six y.
Takes about an hour on 8 cores (the dense gradient needs Sigma^{-1}: about 20 GB at N = 20,000).

Usage:  python3 tests/golden/make_grad_n20000.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402

N, B, H = 20000, 250, 1e-5
FIELDS = [(r, e) for r in (0.1, 0.3, 1.0) for e in (0.01, 0.05)]


def main():
    out = dict(fields=np.array(FIELDS), h=H, ys=[], ll_dense=[], g_dense=[], ll_tlr=[], g_fd=[], quad=[],
               completed=[])
    S = None
    bnd = O.tile_bounds(N, B)
    for r, e in FIELDS:
        th = np.array([1.0, r, e])
        S_f, y = O.simulate_seed0(N, theta=th, seed=0)
        if S is None:
            S = S_f
        assert np.array_equal(S, S_f)
        ll_d, g_d = O.loglik_dense_grad(S, y, th)
        ll, _, info = O.tlr_loglik(S, y, th, bnd, "adaptive", 0.9)
        g = np.full(3, np.nan)
        if ll is not None:
            g[0] = 0.5 * info["quad"] - 0.5 * N
            for k in (1, 2):
                ek = np.zeros(3)
                ek[k] = H
                lp = O.tlr_loglik(S, y, th * np.exp(ek), bnd, "adaptive", 0.9, frozen=info)[0]
                lm = O.tlr_loglik(S, y, th * np.exp(-ek), bnd, "adaptive", 0.9, frozen=info)[0]
                g[k] = (lp - lm) / (2 * H)
        out["ys"].append(y)
        out["ll_dense"].append(ll_d)
        out["g_dense"].append(g_d)
        out["ll_tlr"].append(np.nan if ll is None else ll)
        out["g_fd"].append(g)
        out["quad"].append(info.get("quad", np.nan))
        out["completed"].append(ll is not None)
        print(f"field range={r} eta={e}: dense {ll_d} {g_d}; TLR {ll} {g}", flush=True)
    np.savez_compressed(os.path.join(HERE, "grad_n20000_b250.npz"), S=S, **{k: np.asarray(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
