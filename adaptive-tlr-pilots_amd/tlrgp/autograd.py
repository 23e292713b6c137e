"""torch.autograd bridge: the TLR log-likelihood as a differentiable function of log theta.

forward evaluates log L and its gradient (frozen ranks, see TLRContext.loglik_grad_batch) on the GPU
library; backward scales the stored gradient.  Parameters are log(sigma2, range, eta), the scale on which
the reference's optimizers and samplers work (codes/17_tlr_cholesky_pilot.py:279-281,
codes/38_tlr_cholesky_mcmc.py:105-107).  Non-PD evaluations give log L = -inf and a zero gradient."""
import numpy as np
import torch


class TLRLogLik(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_theta, tctx, rule, value, h):
        lt = log_theta.detach().to("cpu", torch.float64).numpy()
        single = lt.ndim == 1
        th = np.exp(np.atleast_2d(lt))
        r = tctx.loglik_grad_batch(th, rule, value, h, ranks=False)
        ll = np.where(r["completed"], r["loglik"], -np.inf)
        g = np.where(r["completed"][:, None] & np.isfinite(r["grad"]), r["grad"], 0.0)
        ctx.save_for_backward(torch.as_tensor(g[0] if single else g, dtype=log_theta.dtype,
                                              device=log_theta.device))
        out = torch.as_tensor(ll[0] if single else ll, dtype=log_theta.dtype, device=log_theta.device)
        return out

    @staticmethod
    def backward(ctx, g_out):
        (grad,) = ctx.saved_tensors
        if grad.ndim == 1:
            return g_out * grad, None, None, None, None
        return g_out[:, None] * grad, None, None, None, None


def tlr_loglik_torch(log_theta, tctx, rule="adaptive", value=0.9, h=1e-5):
    """log L(exp(log_theta)) with autograd; log_theta is (3,) or (m, 3) (m <= tctx.max_batch)."""
    return TLRLogLik.apply(log_theta, tctx, rule, value, h)
