"""Pin the oracle against the reference's own known answers, closed forms and brute force (CPU)."""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.stats as ss

import matrices as M
from oracle import oracle as orc


def test_identity_known_answer():
    # crates/inla_math/src/solver.rs:106-121
    f = orc.Factor(sp.identity(3, format="csc") * 2.0)
    assert np.array_equal(f.solve(np.array([2.0, 4.0, 6.0])), [1.0, 2.0, 3.0])
    assert np.array_equal(f.takahashi_diag(), [0.5, 0.5, 0.5])
    assert abs(f.log_abs_det() - 3 * math.log(2.0)) < 1e-10


@pytest.mark.parametrize("q,tol", [(M.ar1_q(12, 0.7, 2.5), 1e-9), (M.ar1_q(3, 0.2, 1.0), 1e-10),
                                   (M.besag_ridge_q(M.FOUR_CYCLE_CHORD, 1.5, 0.25), 1e-9),
                                   (M.besag_ridge_q(M.PATH6, 1.0, 0.1), 1e-9)])
def test_reference_takahashi_known_answers(q, tol):
    # sparse_ldlt.rs:613-665: Takahashi vs the dense inverse and vs blocked solves
    f = orc.Factor(q)
    taka = f.takahashi_diag()
    dense = np.diag(np.linalg.inv(q.toarray()))
    by_solves = np.array([f.solve(np.eye(q.shape[0])[i])[i] for i in range(q.shape[0])])
    assert np.max(np.abs(taka - dense)) < tol
    assert np.max(np.abs(taka - by_solves)) < tol
    ld, di = orc.dense_ldlt(q.toarray())
    assert np.max(np.abs(di - dense)) < tol


def test_takahashi_catches_planted_error():
    # a wrong sign in the recurrence must fail the dense comparison
    q = M.ar1_q(12, 0.7, 2.5)
    taka = orc.Factor(q).takahashi_diag()
    dense = np.diag(np.linalg.inv(q.toarray()))
    wrong = taka.copy()
    wrong[3] += 2 * (taka[3] - 1 / np.diag(q.toarray())[3])  # flip the correction term
    assert np.max(np.abs(wrong - dense)) > 1e-3


def test_sparse_matches_dense_tridiagonal():
    # ldlt.rs:522-544
    n = 8
    q = sp.csc_matrix(sp.diags([np.full(n - 1, -0.7), np.full(n, 2.5), np.full(n - 1, -0.7)], [-1, 0, 1]))
    f = orc.Factor(q)
    ld, _ = orc.dense_ldlt(q.toarray())
    assert abs(f.log_abs_det() - ld) < 1e-9
    b = np.arange(1.0, n + 1)
    assert np.max(np.abs(f.solve(b) - np.linalg.solve(q.toarray(), b))) < 1e-9


@pytest.mark.parametrize("n,rho,tau", [(50, 0.5, 1.0), (400, 0.95, 3.0)])
def test_ar1_logdet_closed_form(n, rho, tau):
    f = orc.Factor(M.ar1_q(n, rho, tau))
    exact = n * math.log(tau) + math.log(1 - rho * rho)
    assert abs(f.log_abs_det() - exact) < 1e-10 * abs(exact)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_spd_against_numpy(seed):
    q = M.random_spd(120, 0.05, seed)
    qd = q.toarray()
    for perm in (None, np.arange(120), np.random.default_rng(seed).permutation(120)):
        f = orc.Factor(q, perm)
        assert f.status == 0
        sign, ld = np.linalg.slogdet(qd)
        assert abs(f.log_abs_det() - ld) < 1e-10 * abs(ld)
        b = np.random.default_rng(seed).standard_normal(120)
        x = np.linalg.solve(qd, b)
        assert np.max(np.abs(f.solve(b) - x)) < 1e-10 * np.max(np.abs(x))
        d = np.diag(np.linalg.inv(qd))
        assert np.max(np.abs(f.takahashi_diag() - d)) < 1e-10 * np.max(d)


def test_selinv_entries_on_pattern():
    q = M.random_spd(60, 0.08, 5)
    f = orc.Factor(q)
    inv = np.linalg.inv(q.toarray())
    cols = np.repeat(np.arange(60), np.diff(q.indptr))
    got = f.selinv_entries(q.indices, cols)
    assert np.max(np.abs(got - inv[q.indices, cols])) < 1e-11


def test_indefinite_and_singular_pivots():
    assert orc.Factor(sp.csc_matrix(np.diag([1.0, -3.0, 2.0]))).status == 0
    assert orc.Factor(sp.csc_matrix(np.diag([1.0, 1e-16, 2.0]))).status == 1


# ---------------------------------------------------------------- chooser
def test_time_major_reference_case():
    # ordering.rs:191-204
    q = M.mixture_major(64, 5, 4.0, -0.3, -1.0)
    bw = orc.lower_bandwidth(q)
    assert bw >= 63
    fwd = orc.time_major(q)
    assert fwd is not None
    assert orc.bandwidth_with_fwd(q, fwd) <= 2 * 5
    k, f2 = orc.choose_order(q)
    assert k == 1 and np.array_equal(f2, fwd)


def test_fgn_like_factor_is_sparse():
    # sparse_ldlt.rs:549-570: nnz(L) < n^2 / 8 for n_time=64, n_comp=5
    q = M.mixture_major(64, 5, 6.0, -0.4, -1.2)
    f = orc.Factor(q)
    assert f.nnz_l() < 320 * 320 // 8


def test_rcm_hand_case():
    # path 0-1-2-3-4 relabelled as 2,4,0,3,1: RCM must recover a bandwidth-1 order
    lab = [2, 4, 0, 3, 1]
    rows, cols = [], []
    for a, b in zip(range(4), range(1, 5)):
        rows += [lab[a], lab[b]]
        cols += [lab[b], lab[a]]
    q = sp.csc_matrix((np.ones(8), (rows, cols)), shape=(5, 5)) + sp.identity(5)
    q = sp.csc_matrix(q)
    fwd = orc.rcm(q)
    assert orc.bandwidth_with_fwd(q, fwd) == 1
    # degrees: ends (labels 2 and 1) have degree 1; the lowest (degree, index) start is label 1,
    # BFS from 1: 1, 3, 0, 4, 2; reversed: 2, 4, 0, 3, 1 -> fwd[old] = new
    assert list(fwd) == [2, 4, 0, 3, 1][::-1] or list(np.argsort(fwd)) == [2, 4, 0, 3, 1]


def test_chooser_amd_branch_for_spde():
    import sys
    sys.path.insert(0, "rust-inla_amd")
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(40, 40)
    k, _ = orc.choose_order(g.csc([1, 1, 1]))
    assert k == 0  # lattice bandwidth 2 nx < n/8 and RCM cannot halve it


# ---------------------------------------------------------------- Newton / likelihoods
def test_log_gamma_lanczos():
    for z in (0.3, 1.0, 2.5, 10.0, 123.4, 1e4):
        assert abs(orc.log_gamma(z) - math.lgamma(z)) < 1e-12 * max(1.0, abs(math.lgamma(z)))


def test_likelihoods_against_scipy_and_fd():
    eta = np.linspace(-2, 2, 9)
    h = 1e-6
    cases = [("gaussian", "identity", np.linspace(-1, 1, 9), 4.0, None,
              lambda e, y: ss.norm.logpdf(y, e, 0.5)),
             ("poisson", "log", np.arange(9.0), 1.0, None, lambda e, y: ss.poisson.logpmf(y, np.exp(e))),
             ("binomial", "logit", np.arange(9.0) % 6, 6.0, None,
              lambda e, y: ss.binom.logpmf(y, 6, 1 / (1 + np.exp(-e)))),
             ("nbinomial", "log", np.arange(9.0), 1.0, 5.0,
              lambda e, y: ss.nbinom.logpmf(y, 5, 5 / (5 + np.exp(e))))]
    for fam, link, y, a, b, ref in cases:
        lp, g, hs = orc.eval_lik(fam, link, eta, y, a, b)
        assert np.max(np.abs(lp - ref(eta, y))) < 1e-10
        lpp, _, _ = orc.eval_lik(fam, link, eta + h, y, a, b)
        lpm, _, _ = orc.eval_lik(fam, link, eta - h, y, a, b)
        assert np.max(np.abs(g - (lpp - lpm) / (2 * h))) < 1e-5
        if fam in ("gaussian", "nbinomial"):  # observed Hessian
            _, gp, _ = orc.eval_lik(fam, link, eta + h, y, a, b)
            _, gm, _ = orc.eval_lik(fam, link, eta - h, y, a, b)
            assert np.max(np.abs(hs - (gp - gm) / (2 * h))) < 1e-4


def test_gaussian_mode_is_linear_solve_and_mlik_closed_form():
    # inference.rs:2844-2929: exactly 2 factorisations and 1 solve; mode = linear solve
    rng = np.random.default_rng(0)
    n, m = 30, 45
    q = M.ar1_q(n, 0.6, 2.0)
    A = sp.csc_matrix(rng.random((m, n)) * (rng.random((m, n)) < 0.1))
    y = rng.standard_normal(m)
    prec = 3.0
    r = orc.find_mode(q, "gaussian", "identity", y, prec, A=A)
    assert r.n_factorizations == 2 and r.n_solves == 1
    qpost = q.toarray() + prec * (A.T @ A).toarray()
    x = np.linalg.solve(qpost, prec * A.T @ y)
    assert np.max(np.abs(r.x - x)) < 1e-10 * np.max(np.abs(x))
    S = A.toarray() @ np.linalg.solve(q.toarray() + 1e-12 * np.eye(n), A.T.toarray()) + np.eye(m) / prec
    exact = ss.multivariate_normal(np.zeros(m), S).logpdf(y)
    assert abs(r.mlik - exact) < 1e-8


def test_poisson_mode_satisfies_gradient_condition():
    rng = np.random.default_rng(1)
    n = 40
    q = M.ar1_q(n, 0.8, 1.5)
    y = rng.poisson(2.0, n).astype(float)
    r = orc.find_mode(q, "poisson", "log", y, 1.0)
    assert r.n_factorizations > 2
    g = y - np.exp(r.x) - q @ r.x
    assert np.max(np.abs(g)) < 1e-6


def test_ccd_design_reference_properties():
    # integration.rs:328-347 and 51-59
    for m, npts in ((1, 3), (2, 9), (3, 15), (5, 27)):
        z, w = orc.ccd_design(m, 1.0)
        assert len(z) == npts
        assert abs(w.sum() - 1.0) < 1e-12
    f0, m = 1.3, 3
    z, w = orc.ccd_design(m, f0)
    n_exp = len(z)
    wexp = 1.0 / ((n_exp - 1) * (1.0 + math.exp(-0.5 * f0 * f0) * (f0 * f0 / m - 1.0)))
    assert abs(w[1] - wexp) < 1e-15 and abs(w[0] - (1 - (n_exp - 1) * wexp)) < 1e-15


def test_ordering_reused_for_a_repeated_pattern():
    # the oracle keeps the last pattern's ordering per thread (the reference's
    # factorize_sparse with its thread-local pattern cache, sparse_ldlt.rs:200-207,234):
    # the same pattern with new values gets the same ordering and the results of an
    # explicitly ordered factorisation; another pattern gets its own ordering
    q1 = M.random_spd(300, 0.02, 5)
    q2 = sp.csc_matrix((q1.data * 1.7 + 0.3 * (q1.indices == np.repeat(np.arange(300), np.diff(q1.indptr))),
                        q1.indices, q1.indptr), shape=q1.shape)
    f1 = orc.Factor(q1)
    f2 = orc.Factor(q2)
    assert np.array_equal(f1.perm(), f2.perm())
    g2 = orc.Factor(q2, perm=orc.min_degree(q2))
    assert np.array_equal(f2.perm(), g2.perm())
    assert f2.log_abs_det() == g2.log_abs_det()
    rhs = np.random.default_rng(0).standard_normal(300)
    assert np.array_equal(f2.solve(rhs), g2.solve(rhs))
    q3 = M.random_spd(300, 0.02, 6)
    f3 = orc.Factor(q3)
    assert np.array_equal(f3.perm(), orc.Factor(q3, perm=orc.min_degree(q3)).perm())
    assert abs(f3.log_abs_det() - np.linalg.slogdet(q3.toarray())[1]) < 1e-9 * abs(f3.log_abs_det())
