"""CPU oracle for the adaptive-rank TLR Gaussian-process log-likelihood.

TEST INFRASTRUCTURE ONLY.  This module is the checker: ``tests/``, the smoke entry point and the
CPU-baseline leg of the benchmark may import it; the product path (the HIP library and its Python
facade) never does.

It restates, in NumPy/SciPy, the semantics of the reference scripts

* ``codes/11_tlr_rank_rules.py``  -- kernel, distances, Morton order, tile bounds, dense log-likelihood,
  the data generator recipe;
* ``codes/17_tlr_cholesky_pilot.py`` -- per-tile truncated SVD, QR+SVD recompression, the right-looking
  TLR Cholesky, the tile forward/backward solves and the TLR log-likelihood;
* ``codes/36_likelihood_error_terms.py`` -- the error terms of the assembled approximation (every
  off-diagonal tile truncated by 11's ``truncate_tiles``) against the dense likelihood.

The storage-precision ``mixed`` rule of 17 (adaptive ranks, then every factor pair stored in the lowest
of FP8-E4M3 / FP16 / FP32 / FP64 whose rounding error fits the budget left after truncation, arithmetic
in FP64) is restated in ``round_to`` / ``stored`` / ``round_pair``.  In addition to the reference's
``info`` dictionary the factorization reports the per-tile ranks (initial and final, as dense
lower-triangular nB x nB int arrays), the log-determinant and the quadratic form, which the reference
computes but never writes.

Pinning (see ``tests/test_oracle_golden.py``): the dense log-likelihood, the TLR log-likelihood, stored
values, mean/max factor rank, completion flags and failing tile of every rule are checked against the
reference's committed result JSONs on the reference's own committed inputs; Morton order, tile bounds
and the kernel against closed forms; the rank rule against matrices with known singular values;
``recompress`` against a direct SVD of the product; the TLR factor against the dense Cholesky at a
vanishing tolerance; the rounding of the ``mixed`` rule against closed forms (representable values,
ties to even, saturation, IEEE half/single conversions) and its precision choice by brute force.  The
``mixed`` factorization itself is pinned only statistically (log L within 0.5, tile counts per precision
within 10, stored values within 0.5 % of the reference's run at N = 10k): its rank and precision decisions
feed each other, so implementation details of the FP16/FP8 conversion move single decisions and with
them the rest of the factorization.  The analytic dense gradient is pinned by central finite
differences of ``loglik_dense``.  ``simulate_seed0`` (the benchmark's synthetic field; the reference has no
such generator) is pinned by its own definition: C^{-1} y gives back the PCG64 normal stream drawn after the
sites, which are Morton-sorted, on fresh draws and on the committed N = 10k / 40k fields;
``cholesky_blocked`` against LAPACK's Cholesky, ``cov_full_lowmem`` entry for entry against ``cov_full``,
``tlr_matvec_L`` as the inverse of ``tlr_forward``, ``loglik_dense_parts`` against the reference's dense
quadratic form (``bound_terms_n10000.json#z2``).  ``truncate_tiles`` is pinned on tiles with known singular
values (ranks, tile errors and stored values of both rules), ``bound_terms`` with ``tile_fro`` against every
term the reference's ``bound_terms_n10000.json`` holds for the adaptive rule at 0.9.
"""
import numpy as np
from scipy.linalg import cho_factor, cho_solve, cholesky, solve_triangular, LinAlgError
from scipy.linalg import svd as _scipy_svd

SQRT3 = np.sqrt(3.0)
LOG2PI = np.log(2.0 * np.pi)


# ----------------------------------------------------------------------------- model primitives
def matern32(h, r):
    """Matern nu=3/2 correlation (1 + x) exp(-x), x = sqrt(3) h / r.  Follows 11:35-37."""
    x = SQRT3 * h / r
    return (1.0 + x) * np.exp(-x)


def dist(A, B):
    """Euclidean distances between rows of A (m,2) and B (n,2), sqrt(max(d2, 0)).  Follows 11:56-58."""
    dx = A[:, None, 0] - B[None, :, 0]
    dy = A[:, None, 1] - B[None, :, 1]
    return np.sqrt(np.maximum(dx * dx + dy * dy, 0.0))


def cov_full(S, theta):
    """Dense covariance sigma2 * matern32 + sigma2 * eta * I.  Follows 11:61-65."""
    sigma2, rng_, eta = theta
    K = sigma2 * matern32(dist(S, S), rng_)
    K[np.diag_indices_from(K)] += sigma2 * eta
    return K


def cov_full_lowmem(S, theta, block=2000):
    """The same dense covariance built row block by row block into one N x N array (17:91-103); used
    for N where the broadcasting temporaries of cov_full do not fit in memory."""
    sigma2, rng_, eta = theta
    n = S.shape[0]
    K = np.empty((n, n))
    for i0 in range(0, n, block):
        i1 = min(n, i0 + block)
        K[i0:i1] = sigma2 * matern32(dist(S[i0:i1], S), rng_)
    K[np.diag_indices_from(K)] += sigma2 * eta
    return K


def cholesky_blocked(K, block=4000):
    """In-place lower Cholesky of K by 4000 x 4000 blocks (right-looking); no single BLAS call sees more
    than 2^31 entries, which the bundled OpenBLAS does not survive at N = 40,000.  Returns K (upper part
    zeroed)."""
    from scipy.linalg.blas import dgemm
    n = K.shape[0]
    for k0 in range(0, n, block):
        k1 = min(n, k0 + block)
        K[k0:k1, k0:k1] = cholesky(K[k0:k1, k0:k1], lower=True, check_finite=False)
        if k1 < n:
            P = solve_triangular(K[k0:k1, k0:k1], K[k1:, k0:k1].T, lower=True, check_finite=False).T
            K[k1:, k0:k1] = P
            for i0 in range(k1, n, block):
                i1 = min(n, i0 + block)
                for j0 in range(k1, i1, block):
                    j1 = min(n, j0 + block)
                    K[i0:i1, j0:j1] = dgemm(-1.0, P[i0 - k1:i1 - k1], P[j0 - k1:j1 - k1], 1.0,
                                            K[i0:i1, j0:j1], trans_b=True)
    for k0 in range(0, n, block):
        K[k0:min(n, k0 + block), min(n, k0 + block):] = 0.0
        blk = K[k0:min(n, k0 + block), k0:min(n, k0 + block)]
        blk[np.triu_indices_from(blk, 1)] = 0.0
    return K


def _spread_bits(v):
    """Interleave zeros between the low 32 bits of v (uint64), the usual magic-mask bit spread."""
    v = v & np.uint64(0xFFFFFFFF)
    for shift, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF), (4, 0x0F0F0F0F0F0F0F0F),
                        (2, 0x3333333333333333), (1, 0x5555555555555555)):
        v = (v | (v << np.uint64(shift))) & np.uint64(mask)
    return v


def morton_order(xy, order=10):
    """Stable argsort of the Z-order key on a 2^order grid; x bits even, y bits odd.  Follows 11:40-53."""
    n = 1 << order
    gx = np.clip((xy[:, 0] * n).astype(np.int64), 0, n - 1).astype(np.uint64)
    gy = np.clip((xy[:, 1] * n).astype(np.int64), 0, n - 1).astype(np.uint64)
    key = _spread_bits(gx) | (_spread_bits(gy) << np.uint64(1))
    return np.argsort(key, kind="stable")


def tile_bounds(n, b):
    """[(0,b), (b,2b), ..., (., n)], last tile possibly short.  Follows 11:69-71."""
    edges = list(range(0, n, b)) + [n]
    return [(edges[t], edges[t + 1]) for t in range(len(edges) - 1)]


def loglik_dense(K, y):
    """Exact Gaussian log-likelihood through a dense Cholesky; (None, None) if K is not PD.  Follows 11:112-120."""
    try:
        c, low = cho_factor(K, lower=True, check_finite=False)
    except LinAlgError:
        return None, None
    alpha = cho_solve((c, low), y, check_finite=False)
    logdet = 2.0 * np.log(np.diag(c)).sum()
    return -0.5 * (y @ alpha + logdet + len(y) * LOG2PI), (c, low)


def loglik_dense_parts(S, y, theta):
    """Dense log L with its log-determinant and quadratic form (the split the TLR path also reports)."""
    K = cov_full(S, theta)
    try:
        c, low = cho_factor(K, lower=True, check_finite=False)
    except LinAlgError:
        return None
    z = solve_triangular(c, y, lower=True, check_finite=False)
    logdet = 2.0 * np.log(np.diag(c)).sum()
    quad = float(z @ z)
    return dict(loglik=-0.5 * (quad + logdet + len(y) * LOG2PI), logdet=float(logdet), quad=quad)


def loglik_dense_grad(S, y, theta):
    """Dense log L and its gradient in (log sigma2, log range, log eta).

    dl/dlog(t) = 1/2 a^T K_t a - 1/2 tr(K^{-1} K_t), a = K^{-1} y, with
    K_{log sigma2} = K, K_{log eta} = sigma2 eta I and K_{log range} = sigma2 x^2 exp(-x)
    (x = sqrt(3) h / r: d/dlog r of (1+x)e^{-x} is x^2 e^{-x}).  The mathematics of the model; the
    reference has no gradient."""
    sigma2, r, eta = theta
    n = len(y)
    x = SQRT3 * dist(S, S) / r
    E = np.exp(-x)
    K = sigma2 * (1.0 + x) * E
    K[np.diag_indices_from(K)] += sigma2 * eta
    c, low = cho_factor(K, lower=True, check_finite=False)
    a = cho_solve((c, low), y, check_finite=False)
    Kinv = cho_solve((c, low), np.eye(n), check_finite=False)
    ll = -0.5 * (y @ a + 2.0 * np.log(np.diag(c)).sum() + n * LOG2PI)
    Kr = sigma2 * x * x * E
    g_s = 0.5 * (a @ K @ a) - 0.5 * n
    g_r = 0.5 * (a @ Kr @ a) - 0.5 * np.sum(Kinv * Kr)
    g_e = 0.5 * sigma2 * eta * (a @ a) - 0.5 * sigma2 * eta * np.trace(Kinv)
    return float(ll), np.array([g_s, g_r, g_e])


# ----------------------------------------------------------------------------- data generators
def simulate_reference(n, n_pred=300, theta=(1.0, 0.3, 0.05), seed=20261020):
    """The reference's generator recipe (11:185-200 = 17:313-325 = 38:28-36): PCG64 stream,
    S_all ~ U[0,1]^2, y_all = chol(K_all) z, random permutation into observed/held-out, Morton order
    of the observed sites.  Returns (S, y, Sp, yp)."""
    rng = np.random.default_rng(seed)
    theta = np.asarray(theta, float)
    n_tot = n + n_pred
    S_all = rng.random((n_tot, 2))
    L = np.linalg.cholesky(cov_full(S_all, theta))
    y_all = L @ rng.standard_normal(n_tot)
    del L
    perm = rng.permutation(n_tot)
    obs, hold = perm[:n], perm[n:]
    S, y = S_all[obs], y_all[obs]
    order = morton_order(S)
    return S[order], y[order], S_all[hold], y_all[hold]


def simulate_seed0(n, theta=(1.0, 0.3, 0.05), seed=0):
    """Synthetic field of the benchmark configurations: S ~ U[0,1]^2 from default_rng(seed), Morton
    ordered (stable sort), y = chol(Sigma(theta)) z drawn in dense FP64 on the ordered sites."""
    rng = np.random.default_rng(seed)
    S = rng.random((n, 2))
    S = S[morton_order(S)]
    if n <= 20000:
        L = cholesky(cov_full(S, np.asarray(theta, float)), lower=True, check_finite=False)
    else:
        L = cholesky_blocked(cov_full_lowmem(S, np.asarray(theta, float)))
    z = rng.standard_normal(n)
    y = np.concatenate([L[i0:i0 + 4000] @ z for i0 in range(0, n, 4000)])
    return S, y


# ----------------------------------------------------------------------------- rank rules
def _tail_norms(s):
    """tail[r] = ||(s_r, s_{r+1}, ...)||_2 for r = 0..len(s), tail[len(s)] = 0 (17:113, 17:130)."""
    t = np.sqrt(np.cumsum((s * s)[::-1])[::-1])
    return np.append(t, 0.0)


def _rank_of(s, rule, thr_rel, thr_abs):
    """Smallest r with tail[r] <= thr; thr = eps ||s|| (fixed) or the absolute budget (17:114-115)."""
    tail = _tail_norms(s)
    thr = thr_rel * np.sqrt((s * s).sum()) if rule == "fixed" else thr_abs
    r = int(np.argmax(tail <= thr))
    return r, float(tail[r])


def _svd(A, full):
    try:
        return np.linalg.svd(A, full_matrices=full)
    except np.linalg.LinAlgError:   # gesdd non-convergence: gesvd driver, as 17:111-112 / 17:128-129
        return _scipy_svd(A, full_matrices=full, lapack_driver="gesvd")


def truncate_svd(T, rule, thr_rel, thr_abs, rank=None):
    """Rank-r truncation of a dense tile; T ~ U V^T with U = U_r s_r, V = V_r.  Follows 17:107-116.
    rank: use this rank instead of the rule's (frozen-rank replay; not in the reference)."""
    U, s, Vt = _svd(T, False)
    r, err = _rank_of(s, rule, thr_rel, thr_abs)
    if rank is not None:
        r = min(int(rank), len(s))
        err = float(np.sqrt((s[r:] ** 2).sum()))
    return U[:, :r] * s[:r], Vt[:r].T, err


def recompress(U, V, rule, thr_rel, thr_abs, rank=None):
    """Smallest-rank re-truncation of U V^T via QR(U), QR(V) and the SVD of R_u R_v^T.  Follows 17:119-133.
    rank: as in truncate_svd."""
    if U.shape[1] == 0:
        return U, V, 0.0
    Qu, Ru = np.linalg.qr(U)
    Qv, Rv = np.linalg.qr(V)
    Uw, s, Vwt = _svd(Ru @ Rv.T, True)
    r, err = _rank_of(s, rule, thr_rel, thr_abs)
    if rank is not None:
        r = min(int(rank), len(s))
        err = float(np.sqrt((s[r:] ** 2).sum()))
    return Qu @ (Uw[:, :r] * s[:r]), Qv @ Vwt[:r].T, err


def truncate_tiles(K, bounds, rule, value, m_theta):
    """The assembled approximation Sigma^TLR in place: every off-diagonal tile of K (both triangles) replaced by
    its rank-r truncation by the rule (fixed: relative Frobenius eps = value; adaptive: absolute budget
    value * b * m_theta / N, b the largest tile).  Returns (stored values, max relative tile error, max
    absolute tile error, sqrt(2 sum e_ij^2), ranks in tile order).  Follows 11:74-109."""
    n = K.shape[0]
    b = max(e - s for s, e in bounds)
    stored = sum((e - s) ** 2 for s, e in bounds)
    emax = value * b * m_theta / n
    ranks, max_rel, max_abs, fro2 = [], 0.0, 0.0, 0.0
    for i, (i0, i1) in enumerate(bounds):
        for j, (j0, j1) in enumerate(bounds[:i]):
            U, s_, Vt = _svd(K[i0:i1, j0:j1], False)
            r, err = _rank_of(s_, "fixed" if rule == "fixed" else "adaptive", value, emax)
            Tr = (U[:, :r] * s_[:r]) @ Vt[:r] if r > 0 else np.zeros((i1 - i0, j1 - j0))
            K[i0:i1, j0:j1] = Tr
            K[j0:j1, i0:i1] = Tr.T
            stored += 2 * r * max(i1 - i0, j1 - j0)
            ranks.append(r)
            fro2 += 2 * err ** 2
            max_abs = max(max_abs, err)
            max_rel = max(max_rel, err / max(np.sqrt((s_ ** 2).sum()), 1e-300))
    return stored, max_rel, max_abs, float(np.sqrt(fro2)), ranks


def tile_fro(A, bounds):
    """||A_ij||_F of the lower tiles i > j, keyed (i, j).  As 36:32-38."""
    return {(i, j): float(np.linalg.norm(A[i0:i1, j0:j1]))
            for i, (i0, i1) in enumerate(bounds) for j, (j0, j1) in enumerate(bounds[:i])}


def bound_terms(S, y, theta, bounds, rules):
    """The terms of the assembled approximation's log-likelihood error that 36 decomposes, for each (rule,
    value) in rules: the error ell^TLR - ell (two dense Cholesky factorizations), the first-order term
    T1 = -tr(Sigma^-1 Delta) / 2 + alpha^T Delta alpha / 2 and its tile-wise bounds, the remainder and its
    bounds through E = L^-1 Delta L^-T (||E||_F, ||E||_2, ||E z||^2 = (Delta alpha)^T Sigma^-1 (Delta alpha)),
    and the eta-route bounds.  Same keys as the reference's bound_terms JSON.  Follows 36:75-136 (the
    spectral norm by a dense symmetric eigensolver instead of ARPACK's eigsh)."""
    S = np.asarray(S, float)
    y = np.asarray(y, float)
    theta = np.asarray(theta, float)
    n = len(y)
    b = max(e - s for s, e in bounds)
    m_theta = theta[0] * theta[2]
    K = cov_full(S, theta)
    c, low = cho_factor(K, lower=True, check_finite=False)
    alpha = cho_solve((c, low), y, check_finite=False)
    ll = -0.5 * (y @ alpha + 2.0 * np.log(np.diag(c)).sum() + n * LOG2PI)
    z = solve_triangular(c, y, lower=True, check_finite=False)
    z2 = float(z @ z)
    Kinv = cho_solve((c, low), np.eye(n), check_finite=False)
    L = np.tril(c)
    Kinv_tiles = tile_fro(Kinv, bounds)
    alpha_blk = [float(np.linalg.norm(alpha[i0:i1])) for (i0, i1) in bounds]
    nb = len(bounds)
    sumKinv = sum(Kinv_tiles.values())
    sumalpha = sum(alpha_blk[i] * alpha_blk[j] for i in range(nb) for j in range(i))
    out = dict(theta=theta.tolist(), m_theta=m_theta, dense_loglik=float(ll), z2=z2, sum_Kinv_tile_norms=sumKinv,
               sum_alpha_block_products=sumalpha, rules={})
    for rule, value in rules:
        name = f"{rule}_{value:g}"
        Kt = K.copy()
        stored, max_rel, max_abs, fro, ranks = truncate_tiles(Kt, bounds, rule, value, m_theta)
        emax = value * b * m_theta / n if rule == "adaptive" else None
        llt, _ = loglik_dense(Kt, y)
        if llt is None:
            out["rules"][name] = dict(rule=rule, value=value, completed=False, stored_frac=stored / n ** 2,
                                      max_abs_tile_err=max_abs, sqrt2sum_e2=fro)
            continue
        D = Kt - K
        e_tiles = tile_fro(D, bounds)
        tr_term = float((Kinv * D).sum())
        q_term = float(alpha @ (D @ alpha))
        T1 = -0.5 * tr_term + 0.5 * q_term
        B_ld = sum(Kinv_tiles[k] * e_tiles[k] for k in e_tiles)
        B_q = sum(alpha_blk[i] * alpha_blk[j] * e_tiles[(i, j)] for (i, j) in e_tiles)
        E = solve_triangular(L, D, lower=True, check_finite=False)
        E = solve_triangular(L, E.T, lower=True, check_finite=False)
        E_F = float(np.linalg.norm(E))
        E_2 = float(np.abs(np.linalg.eigvalsh(0.5 * (E + E.T))).max())
        eta_hat = fro / m_theta
        w = D @ alpha
        Ez2 = float(w @ (Kinv @ w))
        err = float(llt - ll)

        def eta_route(eta):
            return None if eta >= 1 else float(0.5 * (n * (-np.log1p(-eta)) + eta / (1 - eta) * z2))
        out["rules"][name] = dict(
            rule=rule, value=value, completed=True, stored_frac=stored / n ** 2, mean_rank=float(np.mean(ranks)),
            max_rank=int(max(ranks)), e_max=emax, max_abs_tile_err=max_abs, max_rel_tile_err=max_rel,
            sqrt2sum_e2=fro, eta_hat=eta_hat, error=err, first_order=T1, first_order_trace_part=-0.5 * tr_term,
            first_order_quadratic_part=0.5 * q_term, remainder=err - T1, T1_bound_tilewise=B_ld + B_q,
            T1_bound_tilewise_logdet=B_ld, T1_bound_tilewise_quadratic=B_q,
            T1_bound_apriori_emax=None if emax is None else emax * (sumKinv + sumalpha), E_fro=E_F, E_2=E_2,
            remainder_bound_logdet=E_F ** 2 / (4 * (1 - E_2)) if E_2 < 1 else None,
            remainder_bound_quadratic=Ez2 / (2 * (1 - E_2)) if E_2 < 1 else None, Ez2=Ez2,
            remainder_bound_quadratic_loose=E_2 ** 2 * z2 / (2 * (1 - E_2)) if E_2 < 1 else None,
            remainder_bound_logdet_apriori=eta_hat ** 2 / (4 * (1 - eta_hat)) if eta_hat < 1 else None,
            eta_route_bound_actual=eta_route(E_2), eta_route_bound_eta_hat=eta_route(eta_hat),
            eta_route_bound_tau=eta_route(value) if rule == "adaptive" else None)
    return out


def rule_thresholds(rule, value, theta, n, b):
    """(thr_rel, thr_abs) of 17:149-150: fixed -> eps; adaptive and mixed -> tau * b * sigma2 * eta / N
    (b = max tile)."""
    sigma2, _, eta = theta
    if rule == "fixed":
        return value, None
    if rule in ("adaptive", "mixed"):
        return None, value * b * sigma2 * eta / n
    raise ValueError(f"unknown rule {rule!r}")


# ----------------------------------------------------------------------------- storage precision (mixed)
PRECISIONS = (8, 16, 32, 64)


def round_to(x, p):
    """Round to the values of a p-bit format and back to float64 (17:41-59): FP32 and FP16 as NumPy's
    conversions; FP8 as 17 defines it: exponent floor(log2|x|) clipped to [-9, 8], 3 fraction bits with
    round-half-even, magnitudes capped at 448 (no subnormal range below the clip)."""
    if p == 64:
        return x
    if p == 32:
        return x.astype(np.float32).astype(np.float64)
    if p == 16:
        return x.astype(np.float16).astype(np.float64)
    a = np.abs(x)
    out = np.zeros_like(x)
    nz = a > 0
    e = np.clip(np.floor(np.log2(a[nz])), -9, 8)
    v = np.round(a[nz] / 2.0 ** e * 8.0) / 8.0 * 2.0 ** e
    v[v > 448.0] = 448.0
    out[nz] = np.sign(x[nz]) * v
    return out


def stored(U, p):
    """Stored values of U in format p (17:62-69): FP8/FP16 scaled so that max|U| maps to half the
    format's largest value (224, 32752), FP32 unscaled."""
    if p == 64 or U.size == 0:
        return U
    sc = np.abs(U).max()
    if sc == 0:
        return U
    cap = {8: 448.0 * 0.5, 16: 65504.0 * 0.5, 32: 1.0}[p]
    scale = cap / sc if p < 32 else 1.0
    return round_to(U * scale, p) / scale


def round_pair(U, V, budget):
    """(U', V', p): the lowest p whose rounding error ||U V^T - U' V'^T||_F is within budget
    (17:72-82); rank 0 stays FP64."""
    if U.shape[1] == 0:
        return U, V, 64
    T = U @ V.T
    for p in PRECISIONS:
        if p == 64:
            return U, V, 64
        Ur, Vr = stored(U, p), stored(V, p)
        if np.linalg.norm(T - Ur @ Vr.T) <= budget:
            return Ur, Vr, p
    return U, V, 64


# ----------------------------------------------------------------------------- TLR Cholesky
def tlr_cholesky(S, theta, bounds, rule, value, frozen=None, progress=None):
    """Right-looking TLR Cholesky of Sigma(theta) without forming it.  Follows 17:137-215.

    Returns (L_diag, L_off, info); (None, None, info) with completed=False, failed_at=k when the
    diagonal tile k is not positive definite.  info adds 'ranks_sigma' and 'ranks_factor' (nB x nB
    int arrays, lower triangle) to the reference's keys, and 'rank_log' ({(k, i, j): rank} of every
    recompression).  frozen: the info of an earlier evaluation whose ranks_sigma and rank_log are replayed
    instead of the rule's choices (frozen ranks, for derivatives; not in the reference).  progress: called
    as progress(phase, done, total) after every tile row of the compression and every factorization step."""
    n = S.shape[0]
    nB = len(bounds)
    b = max(e - s for s, e in bounds)
    sigma2, r, eta = theta
    thr_rel, thr_abs = rule_thresholds(rule, value, theta, n, b)
    mixed = rule == "mixed"   # adaptive ranks + rounded storage (17:146-147, 163-164, 183-184, 203-204)
    if mixed:
        rule = "adaptive"
    prec = np.zeros((nB, nB), np.int64)
    A_diag, A_off = [], {}
    ranks0 = np.zeros((nB, nB), np.int64)
    stored0 = 0
    # initial tiles, row-major over the lower block triangle (17:155-166)
    for i, (i0, i1) in enumerate(bounds):
        D = sigma2 * matern32(dist(S[i0:i1], S[i0:i1]), r)
        D[np.diag_indices_from(D)] += sigma2 * eta
        A_diag.append(D)
        stored0 += (i1 - i0) ** 2
        for j, (j0, j1) in enumerate(bounds[:i]):
            T = sigma2 * matern32(dist(S[i0:i1], S[j0:j1]), r)
            U, V, tail = truncate_svd(T, rule, thr_rel, thr_abs,
                                      None if frozen is None else frozen["ranks_sigma"][i, j])
            if mixed:
                U, V, _ = round_pair(U, V, thr_abs - tail)
            A_off[(i, j)] = (U, V)
            ranks0[i, j] = U.shape[1]
            stored0 += 2 * U.shape[1] * max(i1 - i0, j1 - j0)
        if progress:
            progress("compression", i + 1, nB)
    rank_log = {}
    info = dict(stored_sigma=int(stored0), max_rank_sigma=int(ranks0.max()) if nB > 1 else 0,
                ranks_sigma=ranks0, rank_log=rank_log)
    L_diag, L_off = [None] * nB, {}
    for k in range(nB):
        try:
            Lkk = cholesky(A_diag[k], lower=True, check_finite=False)
        except LinAlgError:
            info.update(completed=False, failed_at=k)
            return None, None, info
        L_diag[k] = Lkk
        # panel: L_ik = U (L_kk^{-1} V)^T (17:179-185)
        for i in range(k + 1, nB):
            U, V = A_off.pop((i, k))
            if U.shape[1]:
                V = solve_triangular(Lkk, V, lower=True, check_finite=False)
            if mixed:
                U, V, prec[i, k] = round_pair(U, V, thr_abs)
            L_off[(i, k)] = (U, V)
        # trailing update A_ij -= L_ik L_jk^T for k < j <= i (17:187-205)
        for i in range(k + 1, nB):
            Ui, Vi = L_off[(i, k)]
            if Ui.shape[1] == 0:
                continue
            for j in range(k + 1, i + 1):
                Uj, Vj = L_off[(j, k)]
                if Uj.shape[1] == 0:
                    continue
                W = Vi.T @ Vj
                if i == j:
                    A_diag[i] -= (Ui @ W) @ Ui.T
                else:
                    Uij, Vij = A_off[(i, j)]
                    Ur, Vr, tail = recompress(np.hstack([Uij, -(Ui @ W)]), np.hstack([Vij, Uj]),
                                              rule, thr_rel, thr_abs,
                                              None if frozen is None else frozen["rank_log"][(k, i, j)])
                    rank_log[(k, i, j)] = Ur.shape[1]
                    if mixed:
                        Ur, Vr, _ = round_pair(Ur, Vr, thr_abs - tail)
                    A_off[(i, j)] = (Ur, Vr)
        if progress:
            progress("factorization", k + 1, nB)
    ranksF = np.zeros((nB, nB), np.int64)
    for (i, j), (U, _) in L_off.items():
        ranksF[i, j] = U.shape[1]
    stored = sum(d.shape[0] ** 2 for d in L_diag) + sum(2 * u.shape[1] * u.shape[0] for u, _ in L_off.values())
    rk = [u.shape[1] for u, _ in L_off.values()]
    info.update(completed=True, stored_factor=int(stored),
                mean_rank_factor=float(np.mean(rk)) if rk else 0.0,
                max_rank_factor=int(max(rk)) if rk else 0, ranks_factor=ranksF)
    if mixed:   # 17:210-214
        byt = sum(d.shape[0] ** 2 * 8 for d in L_diag) + \
            sum(2 * u.shape[1] * u.shape[0] * int(prec[i, j]) // 8 for (i, j), (u, _) in L_off.items())
        info.update(stored_factor_bytes=int(byt), dense_factor_bytes=int(n * (n + 1) // 2 * 8),
                    factor_bytes_frac=byt / (n * (n + 1) / 2 * 8),
                    precision_mix_factor_tiles={p: int(sum(1 for k in L_off if prec[k] == p)) for p in PRECISIONS},
                    precision_factor=prec)
    return L_diag, L_off, info


def tlr_loglik_grad_fd(S, y, theta, bounds, rule, value, h=1e-5):
    """log L and its gradient in (log sigma2, log range, log eta) at frozen ranks: the ranks chosen at theta
    are replayed at theta * exp(+-h e_k) and the central difference is taken (the derivative of a smooth
    function of theta).  Returns (ll, grad, info) or (None, None, info).  Not in the reference (its
    standard errors difference log L with re-chosen ranks, codes/15_pilot_standard_errors.py:21-31)."""
    theta = np.asarray(theta, float)
    ll, _, info = tlr_loglik(S, y, theta, bounds, rule, value)
    if ll is None:
        return None, None, info
    g = np.empty(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        lp = tlr_loglik(S, y, theta * np.exp(e), bounds, rule, value, frozen=info)[0]
        lm = tlr_loglik(S, y, theta * np.exp(-e), bounds, rule, value, frozen=info)[0]
        g[k] = np.nan if lp is None or lm is None else (lp - lm) / (2 * h)
    return ll, g, info


def tlr_forward(L_diag, L_off, bounds, B):
    """Solve L Z = B by tile forward substitution.  Follows 17:230-240."""
    Z = np.zeros_like(B)
    for k, (k0, k1) in enumerate(bounds):
        rhs = B[k0:k1].copy()
        for j, (j0, j1) in enumerate(bounds[:k]):
            U, V = L_off[(k, j)]
            if U.shape[1]:
                rhs -= U @ (V.T @ Z[j0:j1])
        Z[k0:k1] = solve_triangular(L_diag[k], rhs, lower=True, check_finite=False)
    return Z


def tlr_backward(L_diag, L_off, bounds, Z):
    """Solve L^T X = Z by tile backward substitution (L_ij^T = V U^T).  Follows 17:243-256."""
    nB = len(bounds)
    X = np.zeros_like(Z)
    for k in range(nB - 1, -1, -1):
        k0, k1 = bounds[k]
        rhs = Z[k0:k1].copy()
        for i in range(k + 1, nB):
            i0, i1 = bounds[i]
            U, V = L_off[(i, k)]
            if U.shape[1]:
                rhs -= V @ (U.T @ X[i0:i1])
        X[k0:k1] = solve_triangular(L_diag[k], rhs, lower=True, trans="T", check_finite=False)
    return X


def tlr_matvec_L(L_diag, L_off, bounds, Z):
    """Y = L Z for the tile factor.  Follows 17:218-227."""
    Y = np.zeros_like(Z)
    for k, (k0, k1) in enumerate(bounds):
        Y[k0:k1] = L_diag[k] @ Z[k0:k1]
        for j, (j0, j1) in enumerate(bounds[:k]):
            U, V = L_off[(k, j)]
            if U.shape[1]:
                Y[k0:k1] += U @ (V.T @ Z[j0:j1])
    return Y


def tlr_loglik(S, y, theta, bounds, rule, value, frozen=None, progress=None):
    """TLR log-likelihood; (ll, (L_diag, L_off), info) or (None, None, info).  Follows 17:259-266.
    info additionally carries 'logdet' and 'quad' (z^T z); frozen and progress as in tlr_cholesky."""
    L_diag, L_off, info = tlr_cholesky(S, theta, bounds, rule, value, frozen, progress)
    if L_diag is None:
        return None, None, info
    z = tlr_forward(L_diag, L_off, bounds, y[:, None])[:, 0]
    logdet = 2.0 * sum(np.log(np.diag(d)).sum() for d in L_diag)
    quad = float(z @ z)
    info.update(logdet=float(logdet), quad=quad)
    return float(-0.5 * (quad + logdet + len(y) * LOG2PI)), (L_diag, L_off), info


def tlr_predict(S, y, Sp, theta, bounds, fac):
    """Kriging mean and sd at Sp through the TLR solves.  Follows 17:269-276."""
    L_diag, L_off = fac
    Kp = theta[0] * matern32(dist(Sp, S), theta[1])
    alpha = tlr_backward(L_diag, L_off, bounds, tlr_forward(L_diag, L_off, bounds, y[:, None]))[:, 0]
    mu = Kp @ alpha
    W = tlr_forward(L_diag, L_off, bounds, Kp.T)
    var = theta[0] * (1 + theta[2]) - (W ** 2).sum(0)
    return mu, np.sqrt(np.maximum(var, 1e-12))


def crps_gauss(y, mu, sig):
    """Mean CRPS of Gaussian predictive distributions.  Follows 11:142-146."""
    from scipy.special import erf
    z = (y - mu) / sig
    pdf = np.exp(-0.5 * z ** 2) / np.sqrt(2 * np.pi)
    cdf = 0.5 * (1 + erf(z / np.sqrt(2)))
    return (sig * (z * (2 * cdf - 1) + 2 * pdf - 1 / np.sqrt(np.pi))).mean()


def skill(yp, mu, sig):
    """RMSE, CRPS, 90% coverage.  Follows 11:162-165."""
    z = np.abs(yp - mu) / sig
    return dict(rmse=float(np.sqrt(((yp - mu) ** 2).mean())), crps=float(crps_gauss(yp, mu, sig)),
                cover90=float((z <= 1.6448536).mean()))


