"""Extract the reference's per-theta TLR-Cholesky log-likelihoods on its N=5000 field
(codes/37_posterior_oscillation.py; results/posterior_oscillation_{200,300}/posterior_oscillation_n5000.json,
read as text, no code is run): for each theta (theta_true, the posterior mean, 200 or 300 posterior draws) the
dense log L and, per rule (adaptive 0.9, fixed 1e-4), the difference TLR - dense (None where the TLR
factorization failed), written to oscillation_n5000.json.  The field is the reference's generator at N = 5000,
300 held out, seed 20261020 (sim_n5000_obs.npy, made by make_fixtures.py).

Usage:  python3 tests/golden/make_oscillation_goldens.py [/root/reference]"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = {}
    for tag in ("200", "300"):
        with open(os.path.join(ref, "results", f"posterior_oscillation_{tag}", "posterior_oscillation_n5000.json")) as f:
            d = json.load(f)
        out[tag] = dict(b=d["settings"]["b"], thetas=d["thetas"], loglik_exact=d["loglik_exact"],
                        rules={k: v["per_theta"]["cholesky"] for k, v in d["rules"].items()})
    with open(os.path.join(HERE, "oscillation_n5000.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
