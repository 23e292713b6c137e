"""First-round sketch width of the initial compression (TLR_TUNE_SKETCH).  A rank is accepted only with 24
sketch columns to spare, else the tile is redone in the second round at min(96, b) columns, so the width moves
tiles between the rounds: the initial and final ranks must not change, and log L may move only by rounding.
The ranks themselves are pinned to the oracle's full SVD by the parity tests at the default width."""
import os

import numpy as np
import pytest

import tlr_oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

THETAS = np.array([[1.0, 0.3, 0.05], [0.5, 0.6, 0.01], [2.0, 0.1, 0.1], [1.4, 0.44, 0.06]])


@pytest.fixture(scope="module")
def tl():
    import tlrgp
    return tlrgp


@pytest.mark.parametrize("n,b,rule,value", [(2000, 250, "adaptive", 0.9), (10000, 500, "adaptive", 0.9),
                                            (10000, 500, "fixed", 1e-4)])
def test_sketch_width_keeps_ranks(tl, n, b, rule, value):
    if n == 10000:
        a = np.load(os.path.join(GOLDEN, "seed0_n10000.npy"))
        S, y = a[:, :2].copy(), a[:, 2].copy()
    else:
        S, y = O.simulate_seed0(n, seed=0)
    ctx = tl.TLRContext(S, y, b, max_batch=len(THETAS))
    assert ctx.get_tunable("sketch") == 48
    ref = None
    for k in (64, 48, 40, 32):
        ctx.set_tunable("sketch", k)
        assert ctx.get_tunable("sketch") == k
        r = ctx.loglik_batch(THETAS, rule, value)
        print(f"n={n} b={b} {rule} {value} sketch {k}: completed {r['completed'].tolist()}", flush=True)
        if ref is None:
            ref = r
            continue
        np.testing.assert_array_equal(r["completed"], ref["completed"])
        np.testing.assert_array_equal(r["ranks_sigma"], ref["ranks_sigma"])
        ok = ref["completed"]
        np.testing.assert_array_equal(r["ranks_factor"][ok], ref["ranks_factor"][ok])
        np.testing.assert_allclose(r["loglik"][ok], ref["loglik"][ok], rtol=1e-8, atol=0)


def test_sketch_width_rejected_outside_its_range(tl):
    S, y = O.simulate_seed0(2000, seed=0)
    ctx = tl.TLRContext(S, y, 250, max_batch=1)
    for bad in (24, 44, 104, 32.5):
        with pytest.raises(tl.TLRError):
            ctx.set_tunable("sketch", bad)
    assert ctx.get_tunable("sketch") == 48


def test_products_tunable():
    # three products instead of four (one power iteration fewer) keep the ranks at N=2000 and move log L only by
    # the larger subspace defect beside the truncation; the range is enforced
    import tlrgp
    S, y = O.simulate_seed0(2000, seed=0)
    ctx = tlrgp.TLRContext(S, y, 250, max_batch=len(THETAS))
    assert ctx.get_tunable("products") == 4
    print("case products 4 vs 3 N=2000 b=250", flush=True)
    r4 = ctx.loglik_batch(THETAS, "adaptive", 0.9)
    ctx.set_tunable("products", 3)
    r3 = ctx.loglik_batch(THETAS, "adaptive", 0.9)
    np.testing.assert_array_equal(r3["ranks_sigma"], r4["ranks_sigma"])
    np.testing.assert_allclose(r3["loglik"], r4["loglik"], rtol=1e-6)
    for bad in (1, 7, 3.5):
        with pytest.raises(tlrgp.TLRError):
            ctx.set_tunable("products", bad)
    assert ctx.get_tunable("products") == 3
    ctx.close()
