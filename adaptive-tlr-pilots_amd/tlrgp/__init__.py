"""tlrgp: adaptive-rank TLR Gaussian-process log-likelihood on MI355X (HIP, gfx950).

Importing the package loads libtlrgp.so; there is no CPU fallback."""
from ._lib import TLRError, device_info, lib  # noqa: F401  (fails loudly when the library is missing)
from .likelihood import (Lik, TLRContext, TLRFactor, morton_order, neg_loglik_tlr, rank_info,  # noqa: F401
                         tile_bounds, tlr_loglik, tlr_loglik_grad)

__all__ = ["Lik", "TLRContext", "TLRFactor", "TLRError", "device_info", "morton_order", "neg_loglik_tlr",
           "rank_info", "tile_bounds", "tlr_loglik", "tlr_loglik_grad"]
