"""ctypes binding of libtlrgp.so (declarations in include/tlrgp.h).

There is no CPU fallback: if the library is missing, importing this module raises."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtlrgp.so")

TLR_OK, TLR_NOT_PD, TLR_RANK_CAPPED = 0, 1, 2
TLR_ERR_ARG, TLR_ERR_HIP, TLR_ERR_RANKCAP, TLR_ERR_NODEVICE = -1, -2, -3, -4
TLR_RULE_FIXED, TLR_RULE_ADAPTIVE, TLR_RULE_MIXED = 0, 1, 2
RULES_C = {"fixed": TLR_RULE_FIXED, "adaptive": TLR_RULE_ADAPTIVE, "mixed": TLR_RULE_MIXED}
TLR_PREC_FP64, TLR_PREC_FP32, TLR_PREC_BF16 = 0, 1, 2
TLR_TUNE_GUARD, TLR_TUNE_DIAG_WG, TLR_TUNE_GRID_FRAC, TLR_TUNE_SKETCH, TLR_TUNE_PRODUCTS = 0, 1, 2, 3, 4

# every entry point of include/tlrgp.h with its ctypes signature
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p
SIGNATURES = {
    "tlr_ctx_create": (ctypes.c_int, [ctypes.POINTER(_VP), _D, _D, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
    "tlr_ctx_destroy": (ctypes.c_int, [_VP]),
    "tlr_ctx_dims": (ctypes.c_int, [_VP, _I64, _I, _I]),
    "tlr_ctx_set_y": (ctypes.c_int, [_VP, _D]),
    "tlr_loglik": (ctypes.c_int, [_VP, _D, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _D, _I, _I, _I]),
    "tlr_loglik_batch": (ctypes.c_int, [_VP, ctypes.c_int32, _D, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                        _D, _I, _I, _I, _I]),
    "tlr_loglik_batch_graph": (ctypes.c_int, [_VP, ctypes.c_int32, _D, ctypes.c_int32, ctypes.c_double,
                                              ctypes.c_int32, _D, _I, _I, _I, _I]),
    "tlr_loglik_grad_batch": (ctypes.c_int, [_VP, ctypes.c_int32, _D, ctypes.c_int32, ctypes.c_double,
                                             ctypes.c_int32, ctypes.c_double, _D, _D, _I, _I, _I, _I]),
    "tlr_loglik_grad": (ctypes.c_int, [_VP, _D, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                       _D, _D, _I, _I, _I]),
    "tlr_predict": (ctypes.c_int, [_VP, _D, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _D, ctypes.c_int64,
                                   _D, _D, _D]),
    "tlr_solve": (ctypes.c_int, [_VP, _D, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _D, ctypes.c_int64,
                                 ctypes.c_int64, ctypes.c_int32, _D]),
    "tlr_ctx_set_timing": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tlr_ctx_get_timing": (ctypes.c_int, [_VP, _D, _I64, _D, _D]),
    "tlr_ctx_precisions": (ctypes.c_int, [_VP, ctypes.c_int32, _I]),
    "tlr_ctx_arith_counts": (ctypes.c_int, [_VP, ctypes.c_int32, _I64]),
    "tlr_ctx_set_tunable": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_double]),
    "tlr_ctx_stats": (ctypes.c_int, [_VP, _I64]),
    "tlr_ctx_get_tunable": (ctypes.c_int, [_VP, ctypes.c_int32, _D]),
    "tlr_loglik_noalloc": (ctypes.c_int, [_D, _D, ctypes.c_int64, ctypes.c_int32, _D, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int32, _D, _I, _I, _I]),
    "tlr_bound_terms": (ctypes.c_int, [_VP, _D, ctypes.c_int32, _I, _D, _D, _D, _D, _D, _D, _D, _I]),
    "tlr_dense_loglik": (ctypes.c_int, [_VP, ctypes.c_int32, _D, _D]),
    "tlr_last_error": (ctypes.c_char_p, []),
    "tlr_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class TLRError(RuntimeError):
    """An error of the library; code is its negative return code (TLR_ERR_*), None for the facade's own."""

    def __init__(self, message, code=None):
        super().__init__(message)
        self.code = code


TLR_ERR_ARG, TLR_ERR_HIP, TLR_ERR_RANKCAP, TLR_ERR_NODEVICE = -1, -2, -3, -4


def check(rc):
    if rc < 0:
        raise TLRError(f"tlrgp error {rc}: {lib.tlr_last_error().decode()}", rc)
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_I)


def device_info():
    buf = ctypes.create_string_buffer(256)
    check(lib.tlr_device_info(buf, 256))
    return buf.value.decode()


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a
