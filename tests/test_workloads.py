"""Synthetic data generation (host, no GPU): draws of x ~ N(0, Q^-1) have the
covariance Q^-1. Checked on an AR1 chain and on a Besag x AR1 group model against
the dense inverse, with statistical bars of about five standard errors."""

import numpy as np
import scipy.sparse as sp


def _check(Q, draws=3000, seed=0):
    from inla_gpu import workloads as W

    rng = np.random.default_rng(seed)
    X = np.array([W.sample_gmrf(Q, rng) for _ in range(draws)])
    S = np.linalg.inv(Q.toarray())
    emp = X.T @ X / draws
    sd = np.sqrt(np.outer(np.diag(S), np.diag(S)) + S * S) / np.sqrt(draws)  # Wishart entry sd
    return np.max(np.abs(emp - S) / sd)


def test_ar1_draws_have_the_inverse_covariance():
    n, rho = 40, 0.8
    main = np.r_[1.0, np.full(n - 2, 1 + rho * rho), 1.0]
    Q = sp.csc_matrix(sp.diags([np.full(n - 1, -rho), main, np.full(n - 1, -rho)], [-1, 0, 1]))
    assert _check(Q) < 6.0


def test_group_model_draws_have_the_inverse_covariance():
    from inla_gpu import workloads as W

    wl = W.besag_ar1_workload(nr=4, T=5, scale=1.0, sample=False)
    c, _, _, _ = wl.params(wl.theta_true[None])
    Q = sp.csc_matrix(wl.gmrf.csc(c[0]) + 0.5 * sp.identity(wl.n))  # non-singular
    assert _check(Q, draws=3000, seed=1) < 6.0
