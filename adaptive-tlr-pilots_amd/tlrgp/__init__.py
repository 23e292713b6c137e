"""tlrgp: adaptive-rank TLR Gaussian-process log-likelihood on MI355X (HIP, gfx950).

Importing the package loads libtlrgp.so; there is no CPU fallback."""
from ._lib import TLRError, device_info, lib  # noqa: F401  (fails loudly when the library is missing)
from .likelihood import (Lik, TLRContext, TLRFactor, crps_gauss, kl_tlr, morton_order, neg_loglik_tlr,  # noqa: F401
                         rank_info, skill, tile_bounds, tlr_backward, tlr_forward, tlr_loglik, tlr_loglik_grad,
                         tlr_predict)
from . import inference  # noqa: F401,E402  (MLE, MH, HMC and chain diagnostics over the likelihood)
from .certificates import bound_terms  # noqa: F401,E402  (error terms of the assembled approximation)

__all__ = ["Lik", "TLRContext", "TLRFactor", "TLRError", "bound_terms", "crps_gauss", "device_info", "kl_tlr",
           "morton_order", "neg_loglik_tlr", "rank_info", "skill", "tile_bounds", "tlr_backward", "tlr_forward",
           "tlr_loglik", "tlr_loglik_grad", "tlr_predict"]
