"""The autotuner's three searches on a real context (N = 2000, b = 250): launch shapes give identical log L,
the guard search returns one of its guards with the FP32 error inside its margin, the sketch search keeps a
width whose ranks equal the 64-column round's; each leaves its choice set on the context."""
import numpy as np
import pytest

import tlr_oracle as O

pytestmark = pytest.mark.gpu

TH = np.array([[1.0, 0.3, 0.05], [0.5, 0.6, 0.01], [2.0, 0.1, 0.1]])


def test_autotune_searches_on_the_device():
    import tlrgp
    from tlrgp import autotune
    S, y = O.simulate_seed0(2000, seed=0)
    ctx = tlrgp.TLRContext(S, y, 250, max_batch=len(TH))
    quiet = lambda s: print(s, flush=True)   # noqa: E731
    best, table = autotune.tune_launch(ctx, TH, diag_wgs=(6, 9), grid_fracs=(0.75, 1.0), log=quiet)
    assert best["diag_wg"] in (6, 9) and best["grid_frac"] in (0.75, 1.0)
    assert ctx.get_tunable("diag_wg") == best["diag_wg"]
    guard, gtab = autotune.tune_guard(ctx, TH, guards=(1024.0, 16384.0, 1e6), log=quiet)
    assert guard in (1024.0, 16384.0, 1e6) and ctx.get_tunable("guard") == guard
    row = [r for r in gtab if r["guard"] == guard][0]
    assert row["pd_agree"] and row["max_abs_dlogL"] <= 0.5e-2
    k, stab = autotune.tune_sketch(ctx, TH, widths=(64, 48, 40), log=quiet)
    assert k in (64, 48, 40) and ctx.get_tunable("sketch") == k
    assert all(r["same_ranks"] for r in stab), stab
