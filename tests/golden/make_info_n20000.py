"""Dense FP64 observed information on the reference's committed N=20,000 field (results/tlr_cholesky/
sim_n20000_obs.csv) at the reference's rounded dense MLE theta = (1.3603, 0.3364, 0.0361), by the reference's
scheme (codes/15_pilot_standard_errors.py:21-31: second differences of log L in log theta, step 0.02), with
the analytic dense gradient there; written to info_n20000_dense.npz.  The reference's own standard errors
at N = 20,000 (pilot_standard_errors_n20000.json) were computed on data regenerated from the seed, which
this fixture does not assume.  Takes about ten minutes on 8 cores (19 dense Cholesky factorizations).

Usage:  python3 tests/golden/make_info_n20000.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import tlr_oracle as O  # noqa: E402


def main():
    a = np.load(os.path.join(HERE, "sim_n20000_obs.npy"))
    S, y = a[:, :2].copy(), a[:, 2].copy()
    with open(os.path.join(HERE, "standard_errors.json")) as f:
        g = json.load(f)["20000"]
    th = np.array(g["dense"]["theta_hat"])
    h = g["h"]
    x = np.log(th)

    def f(z):
        return O.loglik_dense(O.cov_full(S, np.exp(z)), y)[0]

    p = 3
    H = np.zeros((p, p))
    f0 = f(x)
    for i in range(p):
        ei = np.zeros(p)
        ei[i] = h
        H[i, i] = (f(x + ei) - 2 * f0 + f(x - ei)) / h ** 2
        for j in range(i):
            ej = np.zeros(p)
            ej[j] = h
            H[i, j] = H[j, i] = (f(x + ei + ej) - f(x + ei - ej) - f(x - ei + ej) + f(x - ei - ej)) / (4 * h ** 2)
        print("row", i, H[i], flush=True)
    info = -H
    se_log = np.sqrt(np.diag(np.linalg.inv(info)))
    ll, grad = O.loglik_dense_grad(S, y, th)
    print("info", info, "se_log", se_log, "gradient", grad, flush=True)
    np.savez(os.path.join(HERE, "info_n20000_dense.npz"), theta=th, h=h, loglik=f0, info_log=info, se_log=se_log,
             gradient=grad)


if __name__ == "__main__":
    main()
