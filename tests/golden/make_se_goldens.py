"""Extract the reference's observed-information standard errors (codes/15_pilot_standard_errors.py) from its
committed results (text only, no code is run): results/pilot_standard_errors_n10000.json (dense and
assembled-adaptive fits: theta_hat, info_log, se_log, step h) and pilot_standard_errors_n20000.json (dense:
theta_hat, se_log_theta, the gradient at the rounded MLE, h), written to standard_errors.json.

Usage:  python3 tests/golden/make_se_goldens.py [/root/reference]"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "results", "pilot_standard_errors_n10000.json")) as f:
        a = json.load(f)
    with open(os.path.join(ref, "results", "pilot_standard_errors_n20000.json")) as f:
        b = json.load(f)
    out = {"10000": {"h": a["step"], "b": a["b"],
                     "dense": {k: a["fits"]["dense"][k] for k in ("theta_hat", "loglik", "se_log", "info_log")},
                     "adaptive_0.9_assembled": {k: a["fits"]["adaptive_0.9"][k]
                                                for k in ("theta_hat", "loglik", "se_log", "info_log")}},
           "20000": {"h": b["h"], "dense": {"theta_hat": b["theta_hat"], "se_log": b["se_log_theta"],
                                            "gradient": b["gradient"]}}}
    with open(os.path.join(HERE, "standard_errors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
