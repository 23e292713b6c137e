"""tlr_loglik_noalloc (the allocation-free entry point, its workspace a static device array) against
tlr_loglik on a context, through the C ABI: two calls in a row, the second on a smaller problem, must give
the same log L, log-det, quadratic form, ranks and failure flag bit for bit (the same kernels run the same
per-tile arithmetic; only the persistent grids are smaller).  Sizes beyond the static limits are refused."""
import numpy as np
import pytest

import tlr_oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(lib, S, y, b, theta, rule, value, precision):
    from tlrgp import _lib
    n = len(y)
    nb = (n + b - 1) // b
    S = np.ascontiguousarray(S, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.zeros(3)
    rs = np.zeros((nb, nb), np.int32)
    rf = np.zeros((nb, nb), np.int32)
    fa = np.zeros(1, np.int32)
    rc = lib.tlr_loglik_noalloc(_lib.dptr(S), _lib.dptr(y), n, b, _lib.dptr(th), _lib.RULES_C[rule], float(value),
                                precision, _lib.dptr(out), _lib.iptr(rs), _lib.iptr(rf), _lib.iptr(fa))
    return rc, out, rs, rf, int(fa[0])


def _context(tl, S, y, b, theta, rule, value, precision):
    from tlrgp import _lib
    n = len(y)
    nb = (n + b - 1) // b
    ctx = tl.TLRContext(S, y, b, max_batch=1)
    out = np.zeros(3)
    rs = np.zeros((nb, nb), np.int32)
    rf = np.zeros((nb, nb), np.int32)
    fa = np.zeros(1, np.int32)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    rc = _lib.lib.tlr_loglik(ctx._ptr, _lib.dptr(th), _lib.RULES_C[rule], float(value), precision, _lib.dptr(out),
                             _lib.iptr(rs), _lib.iptr(rf), _lib.iptr(fa))
    ctx.close()
    return rc, out, rs, rf, int(fa[0])


@pytest.mark.parametrize("rule,value,precision", [("adaptive", 0.9, 0), ("fixed", 1e-4, 0), ("adaptive", 0.9, 1)])
def test_noalloc_matches_context_path(rule, value, precision):
    import tlrgp as tl
    from tlrgp import _lib
    for n, b, seed, theta in ((2000, 250, 0, (1.0, 0.3, 0.05)), (1500, 200, 7, (0.8, 0.2, 0.03))):
        S, y = O.simulate_seed0(n, seed=seed)
        print(f"noalloc vs context: N={n} b={b} {rule} {value} precision {precision}", flush=True)
        got = _noalloc(_lib.lib, S, y, b, theta, rule, value, precision)
        want = _context(tl, S, y, b, theta, rule, value, precision)
        assert got[0] == want[0], (got[0], want[0], _lib.lib.tlr_last_error())
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[2], want[2])
        np.testing.assert_array_equal(got[3], want[3])
        assert got[4] == want[4]


def test_noalloc_refuses_large_problems():
    from tlrgp import _lib
    big = np.zeros((4096, 2))
    rc = _lib.lib.tlr_loglik_noalloc(_lib.dptr(big), _lib.dptr(np.zeros(4096)), 4096, 256,
                                     _lib.dptr(np.array([1.0, 0.3, 0.05])), 1, 0.9, 0, _lib.dptr(np.zeros(3)),
                                     None, None, None)
    assert rc == _lib.TLR_ERR_ARG
    assert b"exceed" in _lib.lib.tlr_last_error()
