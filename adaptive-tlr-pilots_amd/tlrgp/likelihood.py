"""Host-side mirror of the reference's likelihood interface, running on the HIP library.

Reference call sites reproduced here (signatures, argument meaning and failure conventions):
  tlr_loglik(S, y, theta, bounds, rule, value) -> (ll | None, fac | None, info)
      codes/17_tlr_cholesky_pilot.py:259-266 (factorization 17:137-215)
  neg_loglik_tlr(logth, S, y, bounds, rule, value, trace) -> -ll | 1e12
      codes/17_tlr_cholesky_pilot.py:279-290
  Lik(S, y, bounds, rule, value)(theta) -> ll | -inf, with .last
      codes/38_tlr_cholesky_mcmc.py:39-56
  tile_bounds(n, b)                     codes/11_tlr_rank_rules.py:69-71
  morton_order(xy, order=10)            codes/11_tlr_rank_rules.py:40-53

Differences: fac is a TLRFactor summary (ranks, logdet, quad) -- the factor itself stays in device
memory; info carries the reference's keys plus ranks_sigma / ranks_factor (nB x nB), logdet and quad.
"""
import hashlib
import time

import numpy as np

from . import _lib
from ._lib import check, dptr, f64, iptr, lib

RULES = {"fixed": _lib.TLR_RULE_FIXED, "adaptive": _lib.TLR_RULE_ADAPTIVE, "mixed": _lib.TLR_RULE_MIXED}
PRECISION_BITS = (8, 16, 32, 64)
PRECISIONS = {"fp64": _lib.TLR_PREC_FP64, "fp32": _lib.TLR_PREC_FP32, "bf16": _lib.TLR_PREC_BF16}


def tile_bounds(n, b):
    """[(0,b), (b,2b), ..., (., n)], the last tile possibly short (11:69-71)."""
    edges = list(range(0, n, b)) + [n]
    return [(edges[t], edges[t + 1]) for t in range(len(edges) - 1)]


def _spread(v):
    v = v & np.uint64(0xFFFFFFFF)
    for sh, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF), (4, 0x0F0F0F0F0F0F0F0F),
                     (2, 0x3333333333333333), (1, 0x5555555555555555)):
        v = (v | (v << np.uint64(sh))) & np.uint64(mask)
    return v


def morton_order(xy, order=10):
    """Stable argsort of the Z-order key on a 2^order grid, x bits even, y bits odd (11:40-53)."""
    n = 1 << order
    gx = np.clip((xy[:, 0] * n).astype(np.int64), 0, n - 1).astype(np.uint64)
    gy = np.clip((xy[:, 1] * n).astype(np.int64), 0, n - 1).astype(np.uint64)
    return np.argsort(_spread(gx) | (_spread(gy) << np.uint64(1)), kind="stable")


def _tile_size_from_bounds(n, bounds):
    """The uniform tile size b of bounds == tile_bounds(n, b); anything else is rejected."""
    if not bounds:
        raise ValueError("empty bounds")
    b = bounds[0][1] - bounds[0][0]
    if list(map(tuple, bounds)) != tile_bounds(n, b):
        raise ValueError("bounds must equal tile_bounds(N, b) (uniform tiles, last one possibly short)")
    return b


class TLRFactor:
    """Summary of a completed device-resident TLR factor: per-tile ranks, log-determinant, quadratic form,
    and the (context, theta, rule) that reproduce it for prediction."""

    def __init__(self, ranks, logdet, quad, ctx=None, theta=None, rule=None, value=None, precision="fp64"):
        self.ranks = ranks
        self.logdet = logdet
        self.quad = quad
        self.ctx, self.theta, self.rule, self.value, self.precision = ctx, theta, rule, value, precision

    def __repr__(self):
        nb = self.ranks.shape[0]
        return f"TLRFactor(nB={nb}, mean_rank={self.ranks[np.tril_indices(nb, -1)].mean():.3f})"


def rank_info(ranks_sigma, ranks_factor, n, b, completed):
    """The reference's info keys from rank arrays: stored_sigma / max_rank_sigma (17:159-167),
    stored_factor / mean_rank_factor / max_rank_factor (17:206-209)."""
    bounds = tile_bounds(n, b)
    nb = len(bounds)
    sizes = np.array([e - s for s, e in bounds])
    il = np.tril_indices(nb, -1)
    diag = int((sizes ** 2).sum())
    rs = ranks_sigma[il]
    maxdim = np.maximum(sizes[il[0]], sizes[il[1]])
    info = dict(stored_sigma=int(diag + (2 * rs * maxdim).sum()), max_rank_sigma=int(rs.max()) if rs.size else 0,
                ranks_sigma=ranks_sigma)
    if completed:
        rf = ranks_factor[il]
        info.update(stored_factor=int(diag + (2 * rf * sizes[il[0]]).sum()),
                    mean_rank_factor=float(rf.mean()) if rf.size else 0.0,
                    max_rank_factor=int(rf.max()) if rf.size else 0, ranks_factor=ranks_factor)
    return info


def mixed_info(prec, ranks_factor, n, b):
    """The reference's byte accounting of the mixed rule (17:210-214): diagonal tiles dense FP64, each factor
    pair 2 r rows_i bits / 8 bytes in its storage precision."""
    bounds = tile_bounds(n, b)
    nb = len(bounds)
    sizes = np.array([e - s for s, e in bounds])
    il = np.tril_indices(nb, -1)
    pr = prec[il]
    byt = int((sizes ** 2).sum() * 8 + (2 * ranks_factor[il] * sizes[il[0]] * pr // 8).sum())
    dense = n * (n + 1) // 2 * 8
    return dict(stored_factor_bytes=byt, dense_factor_bytes=int(dense), factor_bytes_frac=byt / dense,
                precision_mix_factor_tiles={p: int((pr == p).sum()) for p in PRECISION_BITS}, precision_factor=prec)


class TLRContext:
    """Device-resident sites and observations plus workspace for up to max_batch simultaneous theta."""

    def __init__(self, S, y, b, max_batch=1):
        S = f64(S)
        y = f64(y)
        if S.ndim != 2 or S.shape[1] != 2 or y.shape != (S.shape[0],):
            raise ValueError("S must be (N, 2) and y (N,)")
        self.n = S.shape[0]
        self.b = int(b)
        self.max_batch = int(max_batch)
        self._handle = _lib.ctypes.c_void_p()
        check(lib.tlr_ctx_create(_lib.ctypes.byref(self._handle), dptr(S), dptr(y), self.n, self.b, self.max_batch))
        nb = _lib.ctypes.c_int32()
        check(lib.tlr_ctx_dims(self._ptr, None, None, _lib.ctypes.byref(nb)))
        self.nB = nb.value

    @property
    def _ptr(self):
        """The library's context handle; a closed context raises instead of passing a null handle on."""
        h = self.__dict__.get("_handle")
        if not h:
            raise _lib.TLRError("the TLR context is closed")
        return h

    def close(self):
        h = self.__dict__.get("_handle")
        if h:
            lib.tlr_ctx_destroy(h)
            self._handle = _lib.ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, kernel_class):
        """Bracket launches of one kernel class with HIP events (1 Matern products, 2 residual, 3 panel QR,
        5 factorization steps with lookahead, 0 off); resets the accumulators."""
        check(lib.tlr_ctx_set_timing(self._ptr, int(kernel_class)))

    def get_timing(self):
        """(total kernel ms, launches, algorithmic flops, algorithmic HBM bytes) since set_timing."""
        ms = _lib.ctypes.c_double()
        n = _lib.ctypes.c_int64()
        fl = _lib.ctypes.c_double()
        by = _lib.ctypes.c_double()
        check(lib.tlr_ctx_get_timing(self._ptr, _lib.ctypes.byref(ms), _lib.ctypes.byref(n), _lib.ctypes.byref(fl),
                                     _lib.ctypes.byref(by)))
        return ms.value, n.value, fl.value, by.value

    def set_y(self, y):
        check(lib.tlr_ctx_set_y(self._ptr, dptr(f64(y, (self.n,)))))

    def loglik_batch(self, thetas, rule, value, precision="fp64", ranks=True, graph=False):
        """Evaluate m <= max_batch theta in one pass.  Returns dict of arrays: loglik, logdet, quad (NaN when
        not PD), completed (bool), failed_at (-1 when completed), ranks_sigma / ranks_factor (m, nB, nB),
        rank_capped (bool: a tile of that theta needed more than 64 columns, so it has no result; with m > 1 the
        others are evaluated, with m = 1 the call raises).
        graph: replay a HIP graph of this configuration (captured on its first use; same results)."""
        th = f64(np.atleast_2d(thetas))
        m = th.shape[0]
        if th.shape[1] != 3 or m > self.max_batch:
            raise ValueError(f"thetas must be (m, 3) with m <= {self.max_batch}")
        out = np.empty((m, 3))
        st = np.empty(m, np.int32)
        fa = np.empty(m, np.int32)
        rs = np.zeros((m, self.nB, self.nB), np.int32) if ranks else None
        rf = np.zeros((m, self.nB, self.nB), np.int32) if ranks else None
        fn = lib.tlr_loglik_batch_graph if graph else lib.tlr_loglik_batch
        check(fn(self._ptr, m, dptr(th), RULES[rule], float(value), PRECISIONS[precision], dptr(out), iptr(st),
                 iptr(fa), iptr(rs), iptr(rf)))
        res = dict(loglik=out[:, 0], logdet=out[:, 1], quad=out[:, 2], completed=st == _lib.TLR_OK,
                   failed_at=fa, ranks_sigma=rs, ranks_factor=rf, rank_capped=st == _lib.TLR_RANK_CAPPED)
        if rule == "mixed":
            res["precisions"] = np.stack([self.precisions(s) for s in range(m)])
        return res

    TUNABLES = {"guard": _lib.TLR_TUNE_GUARD, "diag_wg": _lib.TLR_TUNE_DIAG_WG, "grid_frac": _lib.TLR_TUNE_GRID_FRAC,
                "sketch": _lib.TLR_TUNE_SKETCH, "products": _lib.TLR_TUNE_PRODUCTS}

    def set_tunable(self, name, value):
        """Set a launch / precision tunable: guard (headroom of the reduced-precision recompressions, in unit
        roundoffs), diag_wg (work groups per tile of the diagonal update), grid_frac (persistent recompression
        grids as a fraction of the resident capacity), sketch (columns of the first compression round, a
        multiple of 8 in [32, min(96, b)]), products (products with T and T^T of a cold compression, 2-6)."""
        check(lib.tlr_ctx_set_tunable(self._ptr, self.TUNABLES[name], float(value)))

    def get_tunable(self, name):
        v = _lib.ctypes.c_double()
        check(lib.tlr_ctx_get_tunable(self._ptr, self.TUNABLES[name], _lib.ctypes.byref(v)))
        return v.value

    def factorizations(self):
        """Factorizations this context has run (a prediction or solve on the resident factor adds none)."""
        v = _lib.ctypes.c_int64()
        check(lib.tlr_ctx_stats(self._ptr, _lib.ctypes.byref(v)))
        return v.value

    def arith_counts(self, slot=0):
        """Recompressions of the last evaluation in batch slot `slot` run in FP64, FP32 and BF16 (the
        reduced-precision modes hand pairs whose budget is too tight on to a higher precision)."""
        c = np.zeros(3, np.int64)
        check(lib.tlr_ctx_arith_counts(self._ptr, int(slot), c.ctypes.data_as(_lib._I64)))
        return dict(fp64=int(c[0]), fp32=int(c[1]), bf16=int(c[2]))

    def precisions(self, slot=0):
        """Storage precision in bits of every factor tile (nB x nB, 0 off the lower triangle) of the last
        evaluation in batch slot `slot` with the mixed rule (PREC_OF of 17:38)."""
        pr = np.zeros((self.nB, self.nB), np.int32)
        check(lib.tlr_ctx_precisions(self._ptr, int(slot), iptr(pr)))
        return pr

    def loglik_grad_batch(self, thetas, rule, value, h=1e-5, precision="fp64", ranks=True):
        """As loglik_batch plus grad (m, 3): d log L / d(log sigma2, log range, log eta) at frozen ranks (the
        sigma2 component exact, the others central differences with step h replaying the recorded ranks).
        Needs max_batch >= 1; the 4 m perturbed evaluations run in chunks of max_batch."""
        th = f64(np.atleast_2d(thetas))
        m = th.shape[0]
        if th.shape[1] != 3 or m > self.max_batch:
            raise ValueError(f"thetas must be (m, 3) with m <= {self.max_batch}")
        out = np.empty((m, 3))
        grad = np.empty((m, 3))
        st = np.empty(m, np.int32)
        fa = np.empty(m, np.int32)
        rs = np.zeros((m, self.nB, self.nB), np.int32) if ranks else None
        rf = np.zeros((m, self.nB, self.nB), np.int32) if ranks else None
        check(lib.tlr_loglik_grad_batch(self._ptr, m, dptr(th), RULES[rule], float(value), PRECISIONS[precision],
                                        float(h), dptr(out), dptr(grad), iptr(st), iptr(fa), iptr(rs), iptr(rf)))
        return dict(loglik=out[:, 0], logdet=out[:, 1], quad=out[:, 2], grad=grad, completed=st == _lib.TLR_OK,
                    failed_at=fa, ranks_sigma=rs, ranks_factor=rf, rank_capped=st == _lib.TLR_RANK_CAPPED)

    def dense_loglik_batch(self, thetas):
        """The exact (dense FP64) log-likelihood at m theta, as loglik_dense of 11:112-120: dict of loglik,
        logdet, quad (NaN when Sigma is not positive definite) and completed.  One N x N device array per call
        (N <= 46340): the reference value beside the TLR evaluation, not its replacement."""
        th = f64(np.atleast_2d(thetas))
        m = th.shape[0]
        if th.shape[1] != 3:
            raise ValueError("thetas must be (m, 3)")
        out = np.empty((m, 3))
        check(lib.tlr_dense_loglik(self._ptr, m, dptr(th), dptr(out)))
        return dict(loglik=out[:, 0], logdet=out[:, 1], quad=out[:, 2], completed=np.isfinite(out[:, 0]))

    def loglik(self, theta, rule, value, precision="fp64"):
        """One theta: (ll | None, TLRFactor | None, info) as tlr_loglik of 17:259-266."""
        r = self.loglik_batch(np.asarray(theta, float)[None, :], rule, value, precision)
        completed = bool(r["completed"][0])
        info = rank_info(r["ranks_sigma"][0], r["ranks_factor"][0], self.n, self.b, completed)
        info["completed"] = completed
        if not completed:
            info["failed_at"] = int(r["failed_at"][0])
            return None, None, info
        info.update(logdet=float(r["logdet"][0]), quad=float(r["quad"][0]))
        if rule == "mixed":
            info.update(mixed_info(r["precisions"][0], r["ranks_factor"][0], self.n, self.b))
        fac = TLRFactor(r["ranks_factor"][0], info["logdet"], info["quad"], self, np.asarray(theta, float).copy(),
                        rule, value, precision)
        return float(r["loglik"][0]), fac, info

    def predict(self, theta, rule, value, Sp, precision="fp64"):
        """Kriging mean and standard deviation at Sp (n_pred, 2) through the TLR factor at theta
        (tlr_predict, 17:269-276); None when the factorization fails."""
        Sp = f64(Sp)
        if Sp.ndim != 2 or Sp.shape[1] != 2:
            raise ValueError("Sp must be (n_pred, 2)")
        mu = np.empty(Sp.shape[0])
        sd = np.empty(Sp.shape[0])
        rc = check(lib.tlr_predict(self._ptr, dptr(f64(theta, (3,))), RULES[rule], float(value), PRECISIONS[precision],
                                   dptr(Sp), Sp.shape[0], dptr(mu), dptr(sd), None))
        if rc == _lib.TLR_NOT_PD:
            return None
        return mu, sd

    def solve(self, theta, rule, value, B, mode="forward", precision="fp64"):
        """Factorize at theta and return L^{-1} B ("forward", tlr_forward 17:230-240), L^{-T} B ("backward",
        tlr_backward 17:243-256) or (L L^T)^{-1} B ("both"); B is (N,) or (N, m).  None when the factorization
        fails.  Second value: (loglik, logdet, quad) of the factorization."""
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        X = np.array(B.reshape(self.n, -1), dtype=np.float64, order="F", copy=True)
        if X.shape[0] != self.n:
            raise ValueError(f"B must have {self.n} rows")
        o3 = np.empty(3)
        rc = check(lib.tlr_solve(self._ptr, dptr(f64(theta, (3,))), RULES[rule], float(value), PRECISIONS[precision],
                                 dptr(X), self.n, X.shape[1], {"forward": 0, "backward": 1, "both": 2}[mode], dptr(o3)))
        if rc == _lib.TLR_NOT_PD:
            return None, o3
        return (X[:, 0] if vec else X), o3


_CTX_CACHE = {}


def _context_for(S, y, b, max_batch=1):
    """Contexts are cached by the content of (S, y) and b: a sampler calling the likelihood repeatedly
    on one field uploads it once.  The cache keeps four; an evicted context is only dropped from it, so it
    lives on (with its device memory) while a factor handle returned by tlr_loglik still refers to it."""
    S = f64(S)
    y = f64(y)
    h = hashlib.blake2b(S.tobytes(), digest_size=16)
    h.update(y.tobytes())
    key = (h.hexdigest(), int(b), S.shape[0])
    ctx = _CTX_CACHE.get(key)
    if ctx is not None and not ctx.__dict__.get("_handle"):   # closed by its user (through a factor handle)
        del _CTX_CACHE[key]
        ctx = None
    if ctx is None or ctx.max_batch < max_batch:
        if len(_CTX_CACHE) >= 4:
            _CTX_CACHE.pop(next(iter(_CTX_CACHE)))
        ctx = TLRContext(S, y, b, max_batch)
        _CTX_CACHE[key] = ctx
    return ctx


def tlr_loglik(S, y, theta, bounds, rule, value, precision="fp64"):
    """TLR log-likelihood on the GPU: (ll | None, fac | None, info).  Drop-in for 17:259-266."""
    S = f64(S)
    b = _tile_size_from_bounds(S.shape[0], bounds)
    return _context_for(S, y, b).loglik(theta, rule, value, precision)


def tlr_loglik_grad(S, y, theta, bounds, rule, value, h=1e-5, precision="fp64"):
    """(ll | None, grad in (log sigma2, log range, log eta) | None, info) at frozen ranks."""
    S = f64(S)
    b = _tile_size_from_bounds(S.shape[0], bounds)
    ctx = _context_for(S, y, b)
    r = ctx.loglik_grad_batch(np.asarray(theta, float)[None, :], rule, value, h, precision)
    completed = bool(r["completed"][0])
    info = rank_info(r["ranks_sigma"][0], r["ranks_factor"][0], ctx.n, ctx.b, completed)
    info["completed"] = completed
    if not completed:
        info["failed_at"] = int(r["failed_at"][0])
        return None, None, info
    info.update(logdet=float(r["logdet"][0]), quad=float(r["quad"][0]))
    return float(r["loglik"][0]), r["grad"][0].copy(), info


def tlr_predict(S, y, Sp, theta, bounds, fac):
    """Kriging mean and sd at Sp with the TLR factor fac returned by tlr_loglik (17:269-276)."""
    if fac is None or fac.ctx is None:
        raise ValueError("fac must be the factor returned by tlr_loglik")
    out = fac.ctx.predict(theta, fac.rule, fac.value, Sp, fac.precision)
    if out is None:
        raise RuntimeError("the factorization failed at theta")
    return out


def crps_gauss(y, mu, sig):
    """Mean CRPS of Gaussian predictive distributions (closed form; 11:142-146)."""
    from scipy.special import erf
    z = (y - mu) / sig
    pdf = np.exp(-0.5 * z ** 2) / np.sqrt(2 * np.pi)
    cdf = 0.5 * (1 + erf(z / np.sqrt(2)))
    return (sig * (z * (2 * cdf - 1) + 2 * pdf - 1 / np.sqrt(np.pi))).mean()


def skill(yp, mu, sig):
    """RMSE, CRPS and 90% interval coverage of held-out predictions (11:162-165)."""
    z = np.abs(yp - mu) / sig
    return dict(rmse=float(np.sqrt(((yp - mu) ** 2).mean())), crps=float(crps_gauss(yp, mu, sig)),
                cover90=float((z <= 1.6448536).mean()))


def tlr_forward(S, y, theta, bounds, rule, value, B):
    """L^{-1} B with the TLR factor of Sigma(theta) (17:230-240, which takes the factor itself)."""
    S = f64(S)
    ctx = _context_for(S, y, _tile_size_from_bounds(S.shape[0], bounds))
    return ctx.solve(theta, rule, value, B, "forward")[0]


def tlr_backward(S, y, theta, bounds, rule, value, Z):
    """L^{-T} Z with the TLR factor of Sigma(theta) (17:243-256)."""
    S = f64(S)
    ctx = _context_for(S, y, _tile_size_from_bounds(S.shape[0], bounds))
    return ctx.solve(theta, rule, value, Z, "backward")[0]


def kl_tlr(S, theta, bounds, rule, value):
    """KL(N(0, Sigma) || N(0, Sigma_TLR)) = (tr(Sigma_TLR^{-1} Sigma) - N + logdet Sigma_TLR - logdet Sigma) / 2
    as codes/31_kl_tlr.py:37-45: the trace is ||L^{-1} C||_F^2 with C the exact dense Cholesky factor of Sigma
    (computed here on the host with SciPy: it is the exact side of the comparison) and L the GPU TLR factor;
    the N right-hand sides go through the GPU forward solve in blocks.  Returns the reference's keys."""
    from scipy.linalg import cholesky
    S = f64(S)
    n = S.shape[0]
    sigma2, r, eta = (float(t) for t in theta)
    d = np.sqrt(np.maximum(((S[:, None, :] - S[None, :, :]) ** 2).sum(-1), 0.0)) if n <= 4000 else None
    if d is None:   # row blocks, one N x N array
        K = np.empty((n, n))
        for i0 in range(0, n, 2000):
            i1 = min(n, i0 + 2000)
            dd = (S[i0:i1, None, 0] - S[None, :, 0]) ** 2 + (S[i0:i1, None, 1] - S[None, :, 1]) ** 2
            x = np.sqrt(np.maximum(dd, 0.0)) * (np.sqrt(3.0) / r)
            K[i0:i1] = sigma2 * ((1.0 + x) * np.exp(-x))
    else:
        x = d * (np.sqrt(3.0) / r)
        K = sigma2 * ((1.0 + x) * np.exp(-x))
    K[np.diag_indices(n)] += sigma2 * eta
    C = cholesky(K, lower=True, overwrite_a=True, check_finite=False)
    del K
    logdet = 2.0 * float(np.log(np.diag(C)).sum())
    ctx = _context_for(S, np.zeros(n), _tile_size_from_bounds(n, bounds))
    X, o3 = ctx.solve(theta, rule, value, C, "forward")
    if X is None:
        return None
    del C
    tr = float((X * X).sum())
    logdet_t = float(o3[1])
    return dict(method=f"{rule} TLR", tau=value, kl=0.5 * (tr - n + logdet_t - logdet), trace=tr, logdet_exact=logdet,
                logdet_approx=logdet_t)


def neg_loglik_tlr(logth, S, y, bounds, rule, value, trace):
    """Objective of the MLE: -ll, or 1e12 when the factorization fails (17:279-290; dense rule excluded).
    A theta with a tile beyond the 64-column rank capacity counts as a failure too (trace["rank_capped"])."""
    theta = np.exp(logth)
    trace["evals"] += 1
    try:
        ll, _, _ = tlr_loglik(S, y, theta, bounds, rule, value)
    except _lib.TLRError as e:
        if e.code != _lib.TLR_ERR_RANKCAP:
            raise
        trace["rank_capped"] = trace.get("rank_capped", 0) + 1
        ll = None
    if ll is None:
        trace["failures"] += 1
        return 1e12
    return -ll


class Lik:
    """Sampler-facing wrapper (38:39-56): lik(theta) -> ll, or -inf when the factorization fails (or a tile
    of theta needs more than the 64-column rank capacity: last["rank_capped"])."""

    def __init__(self, S, y, bounds, rule, value, precision="fp64"):
        self.S, self.y, self.bounds, self.rule, self.value = f64(S), f64(y), bounds, rule, value
        self.precision = precision
        self.last = {}

    def __call__(self, theta):
        t0 = time.time()
        try:
            ll, fac, info = tlr_loglik(self.S, self.y, theta, self.bounds, self.rule, self.value, self.precision)
        except _lib.TLRError as e:
            if e.code != _lib.TLR_ERR_RANKCAP:
                raise
            self.last = dict(ok=False, rank_capped=True, time=time.time() - t0, fac=None,
                             info=dict(completed=False, rank_capped=True))
            return -np.inf
        self.last = dict(ok=ll is not None, time=time.time() - t0, fac=fac, info=info)
        return -np.inf if ll is None else ll
