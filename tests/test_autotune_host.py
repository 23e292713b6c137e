"""Selection logic of tlrgp.autotune on the host, with a scripted stand-in for a context: which guard, launch
shape and sketch width each search keeps, and what it refuses.  No device is needed (the module imports numpy
only; the stand-in's evaluations sleep for the time a setting is meant to take)."""
import importlib.util
import os
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _autotune():
    # load the module file alone: importing the package would load the HIP library
    spec = importlib.util.spec_from_file_location(
        "autotune_alone", os.path.join(ROOT, "adaptive-tlr-pilots_amd", "tlrgp", "autotune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Scripted:
    """Answers loglik_batch from the current tunables: seconds[name][value] of sleep, log L offsets and rank
    changes per setting, FP32 errors per guard."""

    def __init__(self, m=3, nb=4, seconds=None, dlogl=None, rank_change=(), fp32_err=None, fp32_pd_fail=()):
        self.t = {"guard": 16384.0, "diag_wg": 9, "grid_frac": 1.0, "sketch": 48}
        self.m, self.nb = m, nb
        self.seconds = seconds or {}
        self.dlogl = dlogl or {}
        self.rank_change = set(rank_change)
        self.fp32_err = fp32_err or {}
        self.fp32_pd_fail = set(fp32_pd_fail)

    def set_tunable(self, name, value):
        self.t[name] = value

    def get_tunable(self, name):
        return self.t[name]

    def loglik_batch(self, thetas, rule, value, precision="fp64", ranks=True, graph=False):
        m = len(thetas)
        wait = sum(self.seconds.get(k, {}).get(v, 0.0) for k, v in self.t.items())
        time.sleep(wait)
        ll = np.linspace(100.0, 200.0, m)
        for k, v in self.t.items():
            ll = ll * (1.0 + self.dlogl.get(k, {}).get(v, 0.0))
        completed = np.ones(m, bool)
        if precision != "fp64":
            ll = ll + self.fp32_err.get(self.t["guard"], 0.0)
            if self.t["guard"] in self.fp32_pd_fail:
                completed[0] = False
        rs = np.ones((m, self.nb, self.nb), np.int32)
        if ("sketch", self.t["sketch"]) in self.rank_change:
            rs[0, 2, 1] += 1
        return dict(loglik=np.where(completed, ll, np.nan), completed=completed, ranks_sigma=rs,
                    ranks_factor=rs.copy())

    def arith_counts(self, slot=0):
        frac = {256.0: 0.2, 1024.0: 0.1, 16384.0: 0.02}.get(self.t["guard"], 0.0)
        return {"fp64": int(1000 * (1 - frac)), "fp32": int(1000 * frac), "bf16": 0}


TH = np.array([[1.0, 0.3, 0.05], [0.5, 0.6, 0.01], [2.0, 0.1, 0.1]])


def test_sketch_keeps_the_fastest_width_with_equal_ranks():
    at = _autotune()
    ctx = Scripted(seconds={"sketch": {64: 0.06, 56: 0.04, 48: 0.01, 40: 0.002, 32: 0.001}},
                   rank_change=[("sketch", 40)], dlogl={"sketch": {32: 1e-6}})
    chosen, table = at.tune_sketch(ctx, TH, log=lambda s: None)
    assert chosen == 48 and ctx.get_tunable("sketch") == 48
    rows = {r["sketch"]: r for r in table}
    assert not rows[40]["same_ranks"] and not rows[40]["accepted"]            # faster, but ranks moved
    assert rows[32]["same_ranks"] and not rows[32]["accepted"]                # ranks kept, log L moved 1e-6
    assert rows[64]["accepted"] and rows[64]["max_rel_dlogL"] == 0.0


def test_sketch_falls_back_to_the_reference_width():
    at = _autotune()
    ctx = Scripted(seconds={"sketch": {64: 0.02, 48: 0.001}}, rank_change=[("sketch", 48)])
    chosen, _ = at.tune_sketch(ctx, TH, widths=(64, 48), log=lambda s: None)
    assert chosen == 64


def test_guard_takes_the_smallest_within_margin_and_pd():
    at = _autotune()
    ctx = Scripted(fp32_err={256.0: 2e-2, 1024.0: 4e-3, 16384.0: 1e-5, 1e6: 0.0}, fp32_pd_fail=[1024.0])
    chosen, table = at.tune_guard(ctx, TH, guards=(256.0, 1024.0, 16384.0, 1e6), log=lambda s: None)
    # 256: |dlogL| 2e-2 > 0.5 * 1e-2; 1024: within, but a PD flag differs; 16384 is the first that passes
    assert chosen == 16384.0
    rows = {r["guard"]: r for r in table}
    assert not rows[1024.0]["pd_agree"] and rows[16384.0]["pd_agree"]
    assert rows[256.0]["reduced_frac"] > rows[16384.0]["reduced_frac"]


def test_launch_shape_must_not_change_log_l():
    at = _autotune()
    ctx = Scripted(seconds={"diag_wg": {3: 0.02, 6: 0.001, 9: 0.01}, "grid_frac": {0.5: 0.02, 1.0: 0.001}})
    best, _ = at.tune_launch(ctx, TH, diag_wgs=(3, 6, 9), grid_fracs=(0.5, 1.0), log=lambda s: None)
    assert best == {"diag_wg": 6, "grid_frac": 1.0}
    bad = Scripted(dlogl={"diag_wg": {6: 1e-12}})
    with pytest.raises(RuntimeError, match="changed log L"):
        at.tune_launch(bad, TH, diag_wgs=(3, 6), grid_fracs=(1.0,), log=lambda s: None)


def test_apply_sets_every_chosen_tunable():
    at = _autotune()
    ctx = Scripted()
    at.apply(ctx, {"chosen": {"guard": 4096.0, "diag_wg": 12, "grid_frac": 0.75, "sketch": 56}})
    assert ctx.t == {"guard": 4096.0, "diag_wg": 12, "grid_frac": 0.75, "sketch": 56}
