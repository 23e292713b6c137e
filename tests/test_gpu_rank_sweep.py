"""Initial ranks of the GPU compression (the truncation rule of truncate_svd, 17:107-116, reached by a randomized
range finder with an FP64-verified tail) against the reference's joint rank/precision sweep
(codes/18_joint_rank_precision_pilot.py, results/joint_sweep/): on the reference's N=10k sites, 31 settings of the
tile size (250, 500), range (0.05-3) and nugget (0.05-1), the mean and the largest initial rank of the adaptive rule
at 0.9 and of the fixed rule at 1e-4, which the reference computed by a full SVD of every off-diagonal tile.  The
rank sums must be equal (mean x tiles is an integer) and so must the largest ranks."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_initial_ranks_match_joint_sweep():
    import tlrgp
    with open(os.path.join(GOLDEN, "goldens.json")) as f:
        js = json.load(f)["joint_sweep_ranks"]
    a = np.load(os.path.join(GOLDEN, "sim_n10000_obs.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    bad = []
    for b in (250, 500):
        ents = [e for e in js if e["b"] == b]
        th = np.array([e["theta"] for e in ents])
        nb = (10000 + b - 1) // b
        il = np.tril_indices(nb, -1)
        nt = len(il[0])
        ctx = tlrgp.TLRContext(S, y, b, max_batch=len(ents))
        for rule, value in (("adaptive", ents[0]["tau"]), ("fixed", ents[0]["eps"])):
            print(f"case initial ranks N=10000 b={b} {rule}:{value} over {len(ents)} settings", flush=True)
            r = ctx.loglik_batch(th, rule, value)
            for e, rs in zip(ents, r["ranks_sigma"]):
                x = rs[il]
                want_sum, want_max = int(round(e[rule]["mean_rank"] * nt)), e[rule]["max_rank"]
                if int(x.sum()) != want_sum or int(x.max()) != want_max:
                    bad.append((e["run"], rule, int(x.sum()), want_sum, int(x.max()), want_max))
        ctx.close()
    print("mismatches", bad, flush=True)
    assert not bad, bad
