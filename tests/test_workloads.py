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


def test_bym2_poisson_batch_blocks():
    """The heterogeneous BYM2 + Poisson union: each block's Q(theta) is the BYM2
    precision tau ((1 - phi) I + phi Q*) (besag.rs:192-219) of its graph with the
    intercept's prior precision first, A loads the intercept and the region, and the
    coefficient derivatives match central differences."""
    from inla_gpu import workloads as W

    rng = np.random.default_rng(2)
    models = []
    for s in (12, 30, 7):
        adj = W.delaunay_adjacency(s, rng)
        R = W.besag_structure(adj)
        m = W.Bym2Model(adj, R, W.scale_factor_dense(R), rng.poisson(2.0, size=s).astype(float))
        models.append(m)
    wl = W.bym2_poisson_batch(models=models, int_prec=0.01)
    assert wl.k == 3 and wl.n == 12 + 30 + 7 + 3 and list(wl.block_ptr) == [0, 13, 44, 52]
    assert list(wl.obs_ptr) == [0, 12, 42, 49] and wl.A.shape == (49, 52)
    th = np.array([0.3, -0.4])
    for q, m in enumerate(models):
        tau, phi = np.exp(th[0]), 1 / (1 + np.exp(-th[1]))
        ref = np.zeros((m.n + 1, m.n + 1))
        ref[0, 0] = 0.01
        ref[1:, 1:] = tau * ((1 - phi) * np.eye(m.n) + phi * m.scale * m.R.toarray())
        assert np.allclose(wl.block_prior(q, th).toarray(), ref, rtol=1e-14, atol=1e-14)
        Aq = wl.block_A(q).toarray()
        assert np.array_equal(Aq, np.hstack([np.ones((m.n, 1)), np.eye(m.n)]))
        assert np.array_equal(wl.block_y(q), m.y)
    # the union pattern is block diagonal
    cols = np.repeat(np.arange(wl.n), np.diff(wl.indptr))
    blk = np.searchsorted(wl.block_ptr, np.arange(wl.n), side="right") - 1
    assert np.array_equal(blk[cols], blk[wl.indices])
    thetas = rng.normal(size=(2, 3, 2))
    c, dc, _, _ = wl.params(thetas)
    h = 1e-6
    for l in range(2):
        tp, tm = thetas.copy(), thetas.copy()
        tp[..., l] += h
        tm[..., l] -= h
        fd = (wl.params(tp)[0] - wl.params(tm)[0]) / (2 * h)
        assert np.allclose(dc[..., l, :], fd, rtol=1e-7, atol=1e-9)
