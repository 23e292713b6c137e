"""Error terms of the assembled TLR approximation's log-likelihood, on the GPU.

Reference: codes/36_likelihood_error_terms.py:40-139 (the terms) over codes/11_tlr_rank_rules.py:61-65
(Sigma) and 74-109 (truncate_tiles: every off-diagonal tile replaced by its rank-r truncation).  For each rule
at theta, with Delta = Sigma^TLR - Sigma, alpha = Sigma^-1 y, z = L^-1 y (Sigma = L L^T) and
E = L^-1 Delta L^-T: the error ell^TLR - ell, the first-order term T1 = -tr(Sigma^-1 Delta) / 2 +
alpha^T Delta alpha / 2 with its tile-wise bounds, the remainder with its bounds through ||E||_F, ||E||_2 and
||E z||^2 = (Delta alpha)^T Sigma^-1 (Delta alpha), and the eta-route bounds; the same keys as the
reference's bound_terms JSON.

The device work is one library call (tlr_bound_terms): Sigma, the rule's tiles (the initial compression of an
evaluation at theta: the randomized range finder with its FP64-verified rank decision, the reference's ranks)
and Delta from this library's kernels; the dense factorizations, the inverse, the triangular solves for E and
its spectral norm from rocSOLVER / rocBLAS in device memory (five N x N arrays: about 4 GB at N = 10,000).
This module turns the returned quantities into the reference's bounds.

Command line (the reference script's options and output file, on a field given as .npy with columns x, y, z in
Morton order instead of one simulated by the script):
  python -m tlrgp.certificates --field f.npy --b 250 --range 0.3 --eta 0.05 --rules adaptive:0.9,fixed:1e-4 --out dir
"""
import argparse
import json
import math
import os
import time

import numpy as np

from ._lib import check, dptr, f64, iptr, lib
from .likelihood import RULES, _context_for, _tile_size_from_bounds

DEFAULT_RULES = (("adaptive", 0.9), ("adaptive", 0.5), ("fixed", 1e-4), ("fixed", 1e-5))


def bound_terms(S, y, theta, bounds, rules=DEFAULT_RULES):
    """The error terms of 36 for every (rule, value) in rules; returns the reference's dictionary (settings
    aside): theta, m_theta, dense_loglik, z2, sum_Kinv_tile_norms, sum_alpha_block_products and rules[name]
    with name = f"{rule}_{value:g}"."""
    S, y = f64(S), f64(y)
    theta = f64(theta, (3,))
    n = S.shape[0]
    b = _tile_size_from_bounds(n, bounds)
    ctx = _context_for(S, y, b)
    nb = len(bounds)
    tiles = [(i, j) for i in range(nb) for j in range(i)]
    nt = len(tiles)
    sizes = np.array([e - s for s, e in bounds])
    m_theta = float(theta[0] * theta[2])
    nr = len(rules)
    dense = np.zeros(2)
    kinv_t, cov_t = np.zeros(max(nt, 1)), np.zeros(max(nt, 1))
    alpha_b = np.zeros(nb)
    terms = np.zeros((max(nr, 1), 6))
    err = np.zeros((max(nr, 1), max(nt, 1)))
    ranks = np.zeros((max(nr, 1), nb, nb), dtype=np.int32)
    codes = np.array([RULES[r] for r, _ in rules] or [0], dtype=np.int32)
    vals = np.array([float(v) for _, v in rules] or [1.0])
    check(lib.tlr_bound_terms(ctx._ptr, dptr(theta), nr, iptr(codes), dptr(vals), dptr(dense), dptr(kinv_t),
                              dptr(cov_t), dptr(alpha_b), dptr(terms), dptr(err), iptr(ranks)))
    ll, z2 = float(dense[0]), float(dense[1])
    kinv_t, cov_t = kinv_t[:nt], cov_t[:nt]
    pair_ab = np.array([alpha_b[i] * alpha_b[j] for i, j in tiles])
    sumKinv, sumalpha = float(kinv_t.sum()), float(pair_ab.sum())
    out = dict(theta=theta.tolist(), m_theta=m_theta, dense_loglik=ll, z2=z2, sum_Kinv_tile_norms=sumKinv,
               sum_alpha_block_products=sumalpha, rules={})
    for q, (rule, value) in enumerate(rules):
        name = f"{rule}_{value:g}"
        e = err[q, :nt]
        rk = np.array([ranks[q, i, j] for i, j in tiles], dtype=np.int64)
        stored = int((sizes ** 2).sum() + sum(2 * int(r) * max(sizes[i], sizes[j]) for r, (i, j) in zip(rk, tiles)))
        fro = float(np.sqrt(2.0 * (e ** 2).sum()))
        max_abs = float(e.max()) if nt else 0.0
        max_rel = float((e / np.maximum(cov_t, 1e-300)).max()) if nt else 0.0
        llt, tr_term, q_term, E_F, E_2, Ez2 = (float(x) for x in terms[q])
        if not np.isfinite(llt):
            out["rules"][name] = dict(rule=rule, value=value, completed=False, stored_frac=stored / n ** 2,
                                      max_abs_tile_err=max_abs, sqrt2sum_e2=fro)
            continue
        emax = value * b * m_theta / n if rule == "adaptive" else None
        error = llt - ll
        T1 = -0.5 * tr_term + 0.5 * q_term
        B_ld = float((kinv_t * e).sum())
        B_q = float((pair_ab * e).sum())
        eta_hat = fro / m_theta

        def eta_route(eta):
            return None if eta >= 1 else float(0.5 * (n * (-math.log1p(-eta)) + eta / (1 - eta) * z2))
        out["rules"][name] = dict(
            rule=rule, value=value, completed=True, stored_frac=stored / n ** 2, mean_rank=float(rk.mean()),
            max_rank=int(rk.max()), e_max=emax, max_abs_tile_err=max_abs, max_rel_tile_err=max_rel,
            sqrt2sum_e2=fro, eta_hat=eta_hat, error=error, first_order=T1, first_order_trace_part=-0.5 * tr_term,
            first_order_quadratic_part=0.5 * q_term, remainder=error - T1, T1_bound_tilewise=B_ld + B_q,
            T1_bound_tilewise_logdet=B_ld, T1_bound_tilewise_quadratic=B_q,
            T1_bound_apriori_emax=None if emax is None else emax * (sumKinv + sumalpha), E_fro=E_F, E_2=E_2,
            remainder_bound_logdet=E_F ** 2 / (4 * (1 - E_2)) if E_2 < 1 else None,
            remainder_bound_quadratic=Ez2 / (2 * (1 - E_2)) if E_2 < 1 else None, Ez2=Ez2,
            remainder_bound_quadratic_loose=E_2 ** 2 * z2 / (2 * (1 - E_2)) if E_2 < 1 else None,
            remainder_bound_logdet_apriori=eta_hat ** 2 / (4 * (1 - eta_hat)) if eta_hat < 1 else None,
            eta_route_bound_actual=eta_route(E_2), eta_route_bound_eta_hat=eta_route(eta_hat),
            eta_route_bound_tau=eta_route(value) if rule == "adaptive" else None)
    return out


def main(argv=None):
    """Write bound_terms_n{N}.json as the reference script does (36:40-52, 130-136): settings, the dense
    terms and one entry per rule."""
    from .likelihood import tile_bounds
    ap = argparse.ArgumentParser(description="Error terms of the assembled TLR approximation (codes/36)")
    ap.add_argument("--field", required=True, help=".npy with columns x, y, z (Morton-ordered sites, observations)")
    ap.add_argument("--b", type=int, default=250)
    ap.add_argument("--sigma2", type=float, default=1.0)
    ap.add_argument("--range", type=float, default=0.3)
    ap.add_argument("--eta", type=float, default=0.05)
    ap.add_argument("--rules", type=str, default="adaptive:0.9,adaptive:0.5,fixed:1e-4,fixed:1e-5")
    ap.add_argument("--out", type=str, default="bound_terms")
    a = ap.parse_args(argv)
    f = np.load(a.field)
    S, y = f[:, :2].copy(), f[:, 2].copy()
    theta = np.array([a.sigma2, a.range, a.eta])
    rules = []
    for spec in [x for x in a.rules.split(",") if x]:
        rule, value = spec.split(":")
        rules.append((rule, float(value)))
    t0 = time.time()
    out = bound_terms(S, y, theta, tile_bounds(len(y), a.b), rules)
    out = dict(settings=dict(vars(a), n=len(y)), seconds=time.time() - t0, **out)
    for name, r in out["rules"].items():
        if r["completed"]:
            print(f"{name}: stored {r['stored_frac']:.3f} | error {r['error']:+.4f} = T1 {r['first_order']:+.4f} + "
                  f"remainder {r['remainder']:+.4f} | ||E||_2 {r['E_2']:.4f} ||E||_F {r['E_fro']:.4f}", flush=True)
        else:
            print(f"{name}: not positive definite (stored {r['stored_frac']:.3f})", flush=True)
    os.makedirs(a.out, exist_ok=True)
    path = os.path.join(a.out, f"bound_terms_n{len(y)}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("saved " + path, flush=True)
    return out


if __name__ == "__main__":
    main()
