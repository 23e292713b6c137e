"""Callers of the likelihood: maximum likelihood, Metropolis-Hastings and batched Hamiltonian Monte Carlo.

Restated from the reference (semantics, not code):
  fit_mle       Nelder-Mead on neg_loglik_tlr in log theta, x0 = log(theta_true * (1.5, 0.6, 2)),
                xatol=1e-3, fatol=0.05, maxiter=300      codes/17_tlr_cholesky_pilot.py:380-404
  run_chain     random-walk MH in log theta inside the box LO..HI, Robbins-Monro scale adaptation every 50
                iterations toward acceptance 0.3 during burn-in, proposal re-shaped from the chain's own
                covariance at burn/2, checkpoint every ckpt_every iterations with the RNG state
                                                         codes/38_tlr_cholesky_mcmc.py:89-126
  ess           initial-positive-sequence effective sample size      codes/19_tlr_mcmc_gpu.py:157-167
  observed_information  the negative Hessian of log L in log theta and its standard errors, by second
                differences of log L (codes/15_pilot_standard_errors.py:21-31) or by differencing gradients
Added (no reference counterpart): fit_mle_grad (L-BFGS-B with the frozen-rank gradient) and hmc (leapfrog
HMC over many chains, one batched log L + gradient evaluation per leapfrog step).

Checkpoints are plain .npz files; the generator state is stored as JSON text (no pickled objects), and a
resumed chain is bitwise identical to an uninterrupted one."""
import json
import os
import time

import numpy as np

LO = np.log([0.1, 0.02, 0.001])   # box prior in log(sigma2, range, eta), codes/38_tlr_cholesky_mcmc.py:25
HI = np.log([10.0, 2.0, 1.0])


def _autocorr(x):
    """Autocorrelation of a centred 1-d series by a zero-padded FFT (biased estimator, divisor n)."""
    n = len(x)
    f = np.fft.rfft(x, 2 * n)
    acov = np.fft.irfft(f * np.conj(f))[:n] / n
    return acov / acov[0]


def ess(x):
    """Effective sample size of a 1-d chain, Geyer's initial positive sequence over the pairs
    (rho_0 + rho_1), (rho_2 + rho_3), ...: n / (2 sum pairs - 1)  (codes/19_tlr_mcmc_gpu.py:157-167)."""
    x = np.asarray(x, float)
    x = x - x.mean()
    n = len(x)
    if n < 10 or x.var() == 0:
        return float(n)
    rho = _autocorr(x)
    s = 0.0
    for k in range(0, n - 1, 2):
        pair = rho[k] + rho[k + 1]
        if pair <= 0:
            break
        s += pair
    return float(n / max(1e-12, 2 * s - 1))


def ess1(x):
    """Effective sample size with the pairs (rho_1 + rho_2), (rho_3 + rho_4), ...: n / (1 + 2 sum pairs),
    summed while a pair is positive  (codes/38_tlr_cholesky_mcmc.py:129-135)."""
    x = np.asarray(x, float)
    x = x - x.mean()
    n = len(x)
    rho = _autocorr(x)
    tau = 1.0
    t = 1
    while t + 1 < n and rho[t] + rho[t + 1] > 0:
        tau += 2 * (rho[t] + rho[t + 1])
        t += 2
    return n / tau


def split_rhat_ess(chains):
    """Rank-normalised split R-hat and bulk ESS of chains (M, n) (Vehtari et al. 2021), as computed by
    codes/35_mcmc_multichain_diagnostics.py:26-55: halves of every chain form 2M chains; ranks over all
    draws (average ties) map through the normal quantile (r - 3/8) / (size + 1/4); R-hat from the
    within/between variances; ESS from the chain-averaged autocovariance, rho = 1 - (W - acov) / var_hat,
    summed in pairs from lag 0 until a pair is negative, capped at twice the draw count."""
    from scipy import stats
    chains = np.asarray(chains, float)
    M, n = chains.shape
    h = n // 2
    x = np.concatenate([chains[:, :h], chains[:, h:2 * h]], axis=0)
    ranks = stats.rankdata(x, method="average").reshape(x.shape)
    z = stats.norm.ppf((ranks - 0.375) / (x.size + 0.25))
    m, nn = z.shape
    W = z.var(axis=1, ddof=1).mean()
    B = nn * z.mean(axis=1).var(ddof=1)
    var_hat = (nn - 1) / nn * W + B / nn
    rhat = np.sqrt(var_hat / W)
    zc = z - z.mean(axis=1, keepdims=True)
    f = np.fft.rfft(zc, 2 * nn, axis=1)
    acov = (np.fft.irfft(f * np.conj(f), axis=1)[:, :nn] / nn).mean(axis=0)
    rho = 1.0 - (W - acov) / var_hat
    tau = -1.0
    t = 0
    while t + 1 < nn:
        pair = rho[t] + rho[t + 1]
        if pair < 0:
            break
        tau += 2 * pair
        t += 2
    e = m * nn / max(tau, 1e-12)
    return float(rhat), float(min(e, 2 * m * nn))


def pool_chains(chains_log, burn=0):
    """Posterior summary of several chains in log theta (M, iters, 3) after dropping `burn` draws: mean, sd,
    95% interval, per-chain means, split R-hat and bulk ESS per parameter, Monte-Carlo standard error of the
    mean  (the summary of codes/38_tlr_cholesky_mcmc.py:138-160)."""
    th = np.exp(np.asarray(chains_log, float)[:, burn:])
    pooled = th.reshape(-1, 3)
    rh, es = zip(*[split_rhat_ess(th[:, :, k]) for k in range(3)])
    sd = pooled.std(0, ddof=1)
    return dict(n_chains=th.shape[0], mean=pooled.mean(0).tolist(), sd=sd.tolist(),
                ci95=np.percentile(pooled, [2.5, 97.5], axis=0).T.tolist(), chain_means=th.mean(1).tolist(),
                split_rhat=list(rh), bulk_ess=list(es),
                mc_se_of_mean=(sd / np.sqrt(np.maximum(es, 1.0))).tolist())


def fit_mle(S, y, bounds, rule, value, x0, maxiter=300):
    """Nelder-Mead maximum likelihood with the TLR likelihood (ranks re-chosen at every theta)."""
    from scipy import optimize
    from .likelihood import neg_loglik_tlr
    trace = dict(evals=0, failures=0)
    t0 = time.time()
    res = optimize.minimize(neg_loglik_tlr, np.asarray(x0, float), args=(S, y, bounds, rule, value, trace),
                            method="Nelder-Mead", options=dict(xatol=1e-3, fatol=0.05, maxiter=maxiter))
    return dict(theta_hat=np.exp(res.x).tolist(), neg_loglik=float(res.fun), evals=trace["evals"],
                failures=trace["failures"], seconds=time.time() - t0)


def fit_mle_grad(S, y, bounds, rule, value, x0, maxiter=100, h=1e-5):
    """L-BFGS-B in log theta inside the box LO..HI using log L and its frozen-rank gradient.  With the
    adaptive rule the objective re-chooses ranks at every theta (budget proportional to sigma2 * eta); the
    rank changes add a drift to log L that the frozen-rank gradient does not see, so the line search can
    stop short of the optimum a derivative-free search finds (on the same likelihood ridge)."""
    from scipy import optimize
    from .likelihood import tlr_loglik_grad
    trace = dict(evals=0, failures=0)

    def f(x):
        trace["evals"] += 1
        ll, g, _ = tlr_loglik_grad(S, y, np.exp(x), bounds, rule, value, h)
        if ll is None or not np.all(np.isfinite(g)):
            trace["failures"] += 1
            return 1e12, np.zeros(3)
        return -ll, -g

    t0 = time.time()
    res = optimize.minimize(f, np.asarray(x0, float), jac=True, method="L-BFGS-B",
                            bounds=list(zip(LO, HI)), options=dict(maxiter=maxiter))
    return dict(theta_hat=np.exp(res.x).tolist(), neg_loglik=float(res.fun), evals=trace["evals"],
                failures=trace["failures"], seconds=time.time() - t0, success=bool(res.success),
                grad=(-np.asarray(res.jac)).tolist())


def observed_information(S, y, theta, bounds, rule, value, h=0.02, method="gradient", precision="fp64", ctx=None):
    """Observed information (negative Hessian of log L in log theta) at theta, and its inverse's standard
    errors: se_log = sqrt(diag(I^-1)), se_theta = theta * se_log (the delta method of 15:53-55).

    method "gradient": central differences of the GPU gradient, H[:, i] = (g(x + h e_i) - g(x - h e_i)) / 2h,
    symmetrised; the six gradients run as one batch (ranks re-chosen at each point, frozen within each
    gradient).  method "loglik": the reference's scheme, second differences of log L with step h
    (codes/15_pilot_standard_errors.py:21-31), its 19 evaluations in one batch.
    Returns dict(info_log, se_log, se_theta, loglik) or None when an evaluation fails."""
    from .likelihood import TLRContext, _tile_size_from_bounds
    theta = np.asarray(theta, float)
    x = np.log(theta)
    p = 3
    b = _tile_size_from_bounds(len(y), bounds)
    if method == "gradient":
        pts = [x + s * h * np.eye(p)[i] for i in range(p) for s in (1.0, -1.0)]
        c = ctx if ctx is not None and ctx.max_batch >= len(pts) else TLRContext(S, y, b, max_batch=len(pts))
        hg = 1e-5 if precision == "fp64" else 1e-3
        r = c.loglik_grad_batch(np.exp(np.array(pts)), rule, value, h=hg, precision=precision, ranks=False)
        if not r["completed"].all() or not np.isfinite(r["grad"]).all():
            return None
        H = np.empty((p, p))
        for i in range(p):
            H[:, i] = (r["grad"][2 * i] - r["grad"][2 * i + 1]) / (2 * h)
        H = 0.5 * (H + H.T)
        r0 = c.loglik_batch(theta[None, :], rule, value, precision, ranks=False)
        ll = float(r0["loglik"][0])
    elif method == "loglik":
        pts = [x]
        for i in range(p):
            ei = h * np.eye(p)[i]
            pts += [x + ei, x - ei]
            for j in range(i):
                ej = h * np.eye(p)[j]
                pts += [x + ei + ej, x + ei - ej, x - ei + ej, x - ei - ej]
        c = ctx if ctx is not None and ctx.max_batch >= len(pts) else TLRContext(S, y, b, max_batch=len(pts))
        r = c.loglik_batch(np.exp(np.array(pts)), rule, value, precision, ranks=False)
        if not r["completed"].all():
            return None
        f = r["loglik"]
        H = np.empty((p, p))
        q = 1
        f0 = f[0]
        for i in range(p):
            fp, fm = f[q], f[q + 1]
            q += 2
            H[i, i] = (fp - 2 * f0 + fm) / h ** 2
            for j in range(i):
                H[i, j] = H[j, i] = (f[q] - f[q + 1] - f[q + 2] + f[q + 3]) / (4 * h ** 2)
                q += 4
        ll = float(f0)
    else:
        raise ValueError(f"unknown method {method!r}")
    info = -H
    cov = np.linalg.inv(info)
    se_log = np.sqrt(np.diag(cov))
    return dict(info_log=info, se_log=se_log, se_theta=theta * se_log, loglik=ll)


def _save_ckpt(path, **kw):
    tmp = path + ".tmp.npz"
    np.savez(tmp, **kw)
    os.replace(tmp, path)


def run_chain(lik, start, iters, burn, seed, ckpt=None, log=print, ckpt_every=100):
    """Random-walk Metropolis-Hastings in log theta.  lik(theta) returns log L or -inf and sets
    lik.last['ok'].  Returns (chain (iters, 3) log theta, log L per iteration, acceptance, failures, scale,
    proposal factor, seconds)."""
    rng = np.random.default_rng(seed)
    cur = np.array(start, float)
    scale, Lp, shaped = 0.08, np.eye(3), False
    chain = np.zeros((iters, 3))
    lls = np.zeros(iters)
    times = np.zeros(iters)
    acc = acc_win = fail = it0 = 0
    if ckpt and os.path.exists(ckpt):
        ck = np.load(ckpt)
        it0 = int(ck["it"])
        chain[:it0] = ck["chain"][:it0]
        lls[:it0] = ck["lls"][:it0]
        times[:it0] = ck["times"][:it0]
        cur = ck["cur"].copy()
        scale = float(ck["scale"])
        Lp = ck["Lp"].copy()
        shaped = bool(ck["shaped"])
        acc, acc_win, fail = int(ck["acc"]), int(ck["acc_win"]), int(ck["fail"])
        rng.bit_generator.state = json.loads(str(ck["rng_state"]))
        log(f"  resumed at iteration {it0}")
    cur_ll = lik(np.exp(cur))
    if not np.isfinite(cur_ll):
        raise RuntimeError("initial state: factorization failed")
    for it in range(it0, iters):
        t0 = time.time()
        prop = cur + scale * (Lp @ rng.standard_normal(3))
        if np.all(prop > LO) and np.all(prop < HI):
            ll = lik(np.exp(prop))
            if not lik.last["ok"]:
                fail += 1
            elif np.log(rng.random()) < ll - cur_ll:
                cur, cur_ll = prop, ll
                acc += 1
                acc_win += 1
        chain[it] = cur
        lls[it] = cur_ll
        times[it] = time.time() - t0
        if it < burn:
            if (it + 1) % 50 == 0:
                rate = acc_win / 50
                acc_win = 0
                scale *= np.exp(0.5 * (rate - 0.3))
            if (it + 1) == burn // 2 and not shaped:
                cov = np.cov(chain[burn // 4:it + 1].T) + 1e-8 * np.eye(3)
                Lp = np.linalg.cholesky(cov)
                scale = 2.38 / np.sqrt(3)
                shaped = True
        if ckpt and (it + 1) % ckpt_every == 0:
            _save_ckpt(ckpt, it=it + 1, chain=chain, lls=lls, times=times, cur=cur, scale=scale, Lp=Lp, shaped=shaped,
                       acc=acc, acc_win=acc_win, fail=fail, rng_state=np.array(json.dumps(rng.bit_generator.state)))
    return chain, lls, acc / max(iters, 1), fail, scale, Lp, float(times.sum())


def batched_lik(ctx, rule, value, precision="fp64"):
    """thetas (k, 3) -> (log L (k,), ok (k,) bool) through one batched evaluation on a TLRContext
    (k <= ctx.max_batch).  f.rank_capped counts the theta that had no result because a tile needed more than
    the 64-column rank capacity (they count as failed factorizations)."""
    def f(thetas):
        r = ctx.loglik_batch(thetas, rule, value, precision, ranks=False)
        f.rank_capped += int(r["rank_capped"].sum())
        ok = r["completed"].copy()
        return np.where(ok, r["loglik"], -np.inf), ok
    f.rank_capped = 0
    return f


def batched_loglik_grad(ctx, rule, value, h=1e-5, precision="fp64"):
    """log_thetas (k, 3) -> (log L, gradient in log theta) through one batched evaluation (-inf / NaN when
    the factorization fails or a tile needs more than the 64-column rank capacity; f.rank_capped counts the
    latter)."""
    def f(log_thetas):
        r = ctx.loglik_grad_batch(np.exp(log_thetas), rule, value, h, precision, ranks=False)
        f.rank_capped += int(r["rank_capped"].sum())
        ok = r["completed"]
        return np.where(ok, r["loglik"], -np.inf), np.where(ok[:, None], r["grad"], np.nan)
    f.rank_capped = 0
    return f


def concurrent_loglik_grad(parts):
    """One evaluator over chains on several fields: parts is a list of (loglik_grad, k) with each field's
    batched evaluator (batched_loglik_grad on its own TLRContext) and its number of chains; the returned
    function takes the (sum k, 3) log theta of all chains, evaluates every field's share at the same time (one
    host thread per field; each context has its own streams and the library releases the interpreter during
    its calls) and returns the concatenated (log L, gradient)."""
    from concurrent.futures import ThreadPoolExecutor
    offs = np.cumsum([0] + [k for _, k in parts])
    pool = ThreadPoolExecutor(max_workers=len(parts))

    def f(log_thetas):
        futs = [pool.submit(fn, log_thetas[offs[i]:offs[i + 1]]) for i, (fn, _) in enumerate(parts)]
        res = [fu.result() for fu in futs]
        return np.concatenate([r[0] for r in res]), np.concatenate([r[1] for r in res])
    return f


def run_chains_batched(lik_batch, starts, seeds, iters, burn, ckpt=None, ckpt_every=100):
    """M random-walk MH chains advanced in lockstep, the proposals of one iteration evaluated as one batch.
    Chain c follows exactly the steps of run_chain(start=starts[c], seed=seeds[c]): its own generator,
    no evaluation (and no uniform draw) for a proposal outside the box, no uniform draw after a failed
    factorization, the same Robbins-Monro and re-shaping schedule.  lik_batch(thetas (k, 3)) -> (log L, ok).
    Returns a list of M tuples shaped like run_chain's."""
    starts = np.atleast_2d(np.asarray(starts, float))
    M = starts.shape[0]
    rngs = [np.random.default_rng(s) for s in seeds]
    cur = starts.copy()
    scale = np.full(M, 0.08)
    Lp = np.stack([np.eye(3)] * M)
    shaped = np.zeros(M, bool)
    chain = np.zeros((M, iters, 3))
    lls = np.zeros((M, iters))
    acc = np.zeros(M, int)
    acc_win = np.zeros(M, int)
    fail = np.zeros(M, int)
    seconds = 0.0
    it0 = 0
    if ckpt and os.path.exists(ckpt):
        ck = np.load(ckpt)
        it0 = int(ck["it"])
        chain[:, :it0] = ck["chain"][:, :it0]
        lls[:, :it0] = ck["lls"][:, :it0]
        cur, scale, Lp, shaped = ck["cur"].copy(), ck["scale"].copy(), ck["Lp"].copy(), ck["shaped"].copy()
        acc, acc_win, fail = ck["acc"].copy(), ck["acc_win"].copy(), ck["fail"].copy()
        seconds = float(ck["seconds"])
        for g, st in zip(rngs, json.loads(str(ck["rng_state"]))):
            g.bit_generator.state = st
    cur_ll, ok = lik_batch(np.exp(cur))
    if not np.all(ok):
        raise RuntimeError(f"initial state: factorization failed for chains {np.flatnonzero(~ok).tolist()}")
    cur_ll = np.asarray(cur_ll, float).copy()
    for it in range(it0, iters):
        t0 = time.time()
        prop = np.stack([cur[c] + scale[c] * (Lp[c] @ rngs[c].standard_normal(3)) for c in range(M)])
        inside = np.all((prop > LO) & (prop < HI), axis=1)
        idx = np.flatnonzero(inside)
        if idx.size:
            ll, ok = lik_batch(np.exp(prop[idx]))
            for q, c in enumerate(idx):
                if not ok[q]:
                    fail[c] += 1
                elif np.log(rngs[c].random()) < ll[q] - cur_ll[c]:
                    cur[c], cur_ll[c] = prop[c], ll[q]
                    acc[c] += 1
                    acc_win[c] += 1
        chain[:, it] = cur
        lls[:, it] = cur_ll
        seconds += time.time() - t0
        if it < burn:
            if (it + 1) % 50 == 0:
                scale *= np.exp(0.5 * (acc_win / 50 - 0.3))
                acc_win[:] = 0
            if (it + 1) == burn // 2:
                for c in range(M):
                    if not shaped[c]:
                        cov = np.cov(chain[c, burn // 4:it + 1].T) + 1e-8 * np.eye(3)
                        Lp[c] = np.linalg.cholesky(cov)
                        scale[c] = 2.38 / np.sqrt(3)
                        shaped[c] = True
        if ckpt and (it + 1) % ckpt_every == 0:
            _save_ckpt(ckpt, it=it + 1, chain=chain, lls=lls, cur=cur, scale=scale, Lp=Lp, shaped=shaped, acc=acc,
                       acc_win=acc_win, fail=fail, seconds=seconds,
                       rng_state=np.array(json.dumps([g.bit_generator.state for g in rngs])))
    return [(chain[c], lls[c], acc[c] / max(iters, 1), int(fail[c]), float(scale[c]), Lp[c], seconds)
            for c in range(M)]


def hmc(loglik_grad, starts, iters, step=0.01, n_leapfrog=10, seed=0, ckpt=None, ckpt_every=50, log=None):
    """Batched HMC over M chains in log theta with the box prior LO..HI.

    loglik_grad(log_thetas (M, 3)) -> (ll (M,), grad (M, 3)) evaluates every chain at once (ll = -inf when
    the factorization fails; such proposals are rejected).  Each iteration runs n_leapfrog leapfrog steps of
    size `step` (one batched evaluation per step).  Returns dict with the chains (iters, M, 3), log L,
    acceptance per chain, evaluation count and per-leapfrog-step latencies.  log(message), when given, is
    called after every iteration.  ckpt: a file the state is saved to every ckpt_every iterations and resumed
    from (the chains, the generator, the acceptance counts and the latencies of the evaluations so far, so
    that a resumed run reports over all of them)."""
    rng = np.random.default_rng(seed)
    q = np.array(starts, float)
    M = q.shape[0]
    chain = np.zeros((iters, M, 3))
    lls = np.zeros((iters, M))
    acc = np.zeros(M, int)
    lat = []
    it0 = 0
    if ckpt and os.path.exists(ckpt):
        ck = np.load(ckpt)
        it0 = int(ck["it"])
        chain[:it0] = ck["chain"][:it0]
        lls[:it0] = ck["lls"][:it0]
        q = ck["q"].copy()
        acc = ck["acc"].copy()
        rng.bit_generator.state = json.loads(str(ck["rng_state"]))
        if "lat" in ck.files:
            lat = [float(t) for t in ck["lat"]]

    def evaluate(x):
        t0 = time.perf_counter()
        inside = np.all((x > LO) & (x < HI), axis=1)
        ll, g = loglik_grad(np.clip(x, LO, HI))
        lat.append(time.perf_counter() - t0)
        ll = np.where(inside & np.isfinite(ll), ll, -np.inf)
        g = np.where(np.isfinite(g), g, 0.0)
        return ll, g

    ll_q, g_q = evaluate(q)
    for it in range(it0, iters):
        p0 = rng.standard_normal((M, 3))
        x, p, g = q.copy(), p0 + 0.5 * step * g_q, None
        ll_x = ll_q
        for s in range(n_leapfrog):
            x = x + step * p
            ll_x, g = evaluate(x)
            p = p + (step if s < n_leapfrog - 1 else 0.5 * step) * g
        h_old = -ll_q + 0.5 * (p0 * p0).sum(1)
        h_new = -ll_x + 0.5 * (p * p).sum(1)
        u = np.log(rng.random(M))
        ok = np.isfinite(h_new) & (u < h_old - h_new)
        q[ok] = x[ok]
        ll_q = np.where(ok, ll_x, ll_q)
        g_q[ok] = g[ok]
        acc += ok
        chain[it] = q
        lls[it] = ll_q
        if ckpt and (it + 1) % ckpt_every == 0:
            _save_ckpt(ckpt, it=it + 1, chain=chain, lls=lls, q=q, acc=acc,
                       rng_state=np.array(json.dumps(rng.bit_generator.state)), lat=np.array(lat))
        if log:
            log(f"iteration {it + 1}/{iters}: {len(lat)} batched evaluations, last leapfrog step {lat[-1]:.3f} s, "
                f"accepted {int(ok.sum())}/{M}")
    lat = np.array(lat)
    # acc holds the acceptances of every iteration, the resumed ones included
    return dict(chain=chain, lls=lls, acceptance=(acc / max(iters, 1)).tolist(), evaluations=len(lat) * M,
                step_latency_p50=float(np.percentile(lat, 50)) if lat.size else None,
                step_latency_p99=float(np.percentile(lat, 99)) if lat.size else None,
                seconds=float(lat.sum()))
