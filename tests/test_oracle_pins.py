"""Pins of every exported oracle function by what fixes it independently of the
oracle's own code: the reference's known answers, closed forms, brute force on tiny
inputs and mathematical identities. Each pin is then re-run with a deliberate
mistake switched on inside the oracle (oracle.planted): a pin that still passes
with its function broken would pin nothing, so test_planted_mistake_is_caught
requires every listed pin to fail for its mistake (CPU only).
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.stats as ss

import matrices as M
from oracle import oracle as orc

KAPPA = 3.059023205018258e6  # constraints.rs:268


# ------------------------------------------------------------------ ordering
def pin_bandwidths():
    n = 10
    tri = sp.csc_matrix(sp.diags([np.ones(n - 1), np.full(n, 2.0), np.ones(n - 1)], [-1, 0, 1]))
    assert orc.lower_bandwidth(tri) == 1
    far = sp.csc_matrix(tri + sp.csc_matrix(([1.0, 1.0], ([9, 0], [0, 9])), shape=(n, n)))
    assert orc.lower_bandwidth(far) == 9
    assert orc.bandwidth_with_fwd(far, np.arange(n)) == 9
    # reversing the labels keeps the bandwidth; a relabelling that puts 0 and 9
    # next to each other on a path 0-9 measures the new distances
    assert orc.bandwidth_with_fwd(tri, np.arange(n)[::-1].copy()) == 1
    fwd = np.array([0, 2, 4, 6, 8, 9, 7, 5, 3, 1])
    assert orc.bandwidth_with_fwd(tri, fwd) == 2


def _relabelled_path():
    lab = [2, 4, 0, 3, 1]  # path 0-1-2-3-4 under the labels lab
    rows, cols = [], []
    for a, b in zip(range(4), range(1, 5)):
        rows += [lab[a], lab[b]]
        cols += [lab[b], lab[a]]
    return sp.csc_matrix(sp.csc_matrix((np.ones(8), (rows, cols)), shape=(5, 5)) + sp.identity(5))


def pin_rcm():
    q = _relabelled_path()
    fwd = orc.rcm(q)
    assert sorted(fwd) == list(range(5))
    assert orc.bandwidth_with_fwd(q, fwd) == 1
    # the BFS starts at the lowest-index vertex of minimum degree (label 1, an end of
    # the path) and the order is reversed: that vertex is numbered last
    assert fwd[1] == 4


def pin_time_major():
    # ordering.rs:191-204
    q = M.mixture_major(64, 5, 4.0, -0.3, -1.0)
    assert orc.lower_bandwidth(q) >= 63
    fwd = orc.time_major(q)
    assert fwd is not None and orc.bandwidth_with_fwd(q, fwd) <= 2 * 5
    k, f2 = orc.choose_order(q)
    assert k == 1 and np.array_equal(f2, fwd)


def pin_chooser():
    import sys

    sys.path.insert(0, "rust-inla_amd")
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(40, 40)
    k, _ = orc.choose_order(g.csc([1, 1, 1]))
    assert k == 0  # lattice bandwidth 2 nx < n/8 and RCM cannot halve it
    # a shuffled band: RCM recovers it (halves the bandwidth) and is chosen
    n = 400
    band = sp.diags([np.ones(n - 3), np.ones(n - 1), np.full(n, 4.0), np.ones(n - 1), np.ones(n - 3)],
                    [-3, -1, 0, 1, 3])
    p = np.random.default_rng(0).permutation(n)
    q = sp.csc_matrix(band.tocsr()[p][:, p])
    k, fwd = orc.choose_order(q)
    assert k == 2 and orc.bandwidth_with_fwd(q, fwd) * 2 < orc.lower_bandwidth(q)


def _elimination_check(q, order):
    n = q.shape[0]
    adj = [set() for _ in range(n)]
    qc = sp.coo_matrix(q)
    for r, c in zip(qc.row, qc.col):
        if r != c:
            adj[r].add(c)
            adj[c].add(r)
    alive = set(range(n))
    for v in order:
        dmin = min(len(adj[u]) for u in alive)
        assert len(adj[v]) == dmin, (v, len(adj[v]), dmin)
        nb = adj[v]
        for u in nb:
            adj[u] |= nb
            adj[u].discard(u)
            adj[u].discard(v)
        alive.remove(v)
        adj[v] = set()


def pin_min_degree():
    # each step eliminates a vertex of minimum degree in the current elimination graph
    star = sp.lil_matrix((9, 9))
    for i in range(1, 9):
        star[0, i] = star[i, 0] = 1.0
    star.setdiag(4.0)
    for q in [sp.csc_matrix(star), M.random_spd(30, 0.12, 3), M.random_spd(40, 0.06, 4)]:
        order = orc.min_degree(q)
        assert sorted(order) == list(range(q.shape[0]))
        _elimination_check(q, order)


# ------------------------------------------------------------------ factor
def _symbolic_nnz_l(q, perm):
    """Structural nnz(L) (with the diagonal) of P Q P^T by explicit elimination."""
    b = (sp.csc_matrix(q)[perm][:, perm].toarray() != 0)
    n = b.shape[0]
    low = np.tril(b, -1)
    for k in range(n):
        rows = np.nonzero(low[k + 1:, k])[0] + k + 1
        for i in rows:  # eliminating k connects its lower neighbours
            low[rows[rows > i], i] = True
    return int(low.sum()) + n


def pin_factor_brute():
    for seed in (0, 1):
        q = M.random_spd(40, 0.08, seed)
        qd = q.toarray()
        rng = np.random.default_rng(seed)
        for perm in (None, rng.permutation(40)):
            f = orc.Factor(q, perm)
            assert f.status == 0
            p = f.perm()
            assert sorted(p) == list(range(40))
            if perm is not None:
                assert np.array_equal(p, perm)
            mp = qd[np.ix_(p, p)]
            d_ref = np.diag(np.linalg.cholesky(mp)) ** 2
            assert np.max(np.abs(f.d() - d_ref)) < 1e-12 * np.max(d_ref)
            assert f.nnz_l() == _symbolic_nnz_l(q, p)
            _, ld = np.linalg.slogdet(qd)
            assert abs(f.log_abs_det() - ld) < 1e-11 * abs(ld)
            b = rng.standard_normal(40)
            x = np.linalg.solve(qd, b)
            assert np.max(np.abs(f.solve(b) - x)) < 1e-11 * np.max(np.abs(x))
            di = np.diag(np.linalg.inv(qd))
            assert np.max(np.abs(f.takahashi_diag() - di)) < 1e-11 * np.max(di)


def pin_logdet_indefinite():
    f = orc.Factor(sp.csc_matrix(np.diag([1.0, -3.0, 2.0])))
    assert f.status == 0
    assert abs(f.log_abs_det() - math.log(6.0)) < 1e-14


def pin_takahashi_known():
    # sparse_ldlt.rs:613-665
    for q, tol in [(M.ar1_q(12, 0.7, 2.5), 1e-9), (M.ar1_q(3, 0.2, 1.0), 1e-10),
                   (M.besag_ridge_q(M.FOUR_CYCLE_CHORD, 1.5, 0.25), 1e-9),
                   (M.besag_ridge_q(M.PATH6, 1.0, 0.1), 1e-9)]:
        taka = orc.Factor(q).takahashi_diag()
        assert np.max(np.abs(taka - np.diag(np.linalg.inv(q.toarray())))) < tol


def pin_selinv_entries():
    q = M.random_spd(60, 0.08, 5)
    f = orc.Factor(q)
    inv = np.linalg.inv(q.toarray())
    cols = np.repeat(np.arange(60), np.diff(q.indptr))
    assert np.max(np.abs(f.selinv_entries(q.indices, cols) - inv[q.indices, cols])) < 1e-11


def pin_dense_ldlt():
    q = M.random_spd(30, 0.2, 7).toarray()
    ld, di = orc.dense_ldlt(q)
    assert abs(ld - np.linalg.slogdet(q)[1]) < 1e-11 * abs(ld)
    assert np.max(np.abs(di - np.diag(np.linalg.inv(q)))) < 1e-11 * np.max(di)


def pin_lgamma():
    for z in (0.3, 1.0, 2.5, 10.0, 123.4, 1e4):
        assert abs(orc.log_gamma(z) - math.lgamma(z)) < 1e-12 * max(1.0, abs(math.lgamma(z)))


# ------------------------------------------------------------------ likelihoods
def pin_likelihoods():
    """log-densities against scipy; gradients and Newton weights against central
    differences. The Poisson-log and binomial-logit weights are Fisher information,
    which for these canonical links equals the observed Hessian, so the derivative
    of the gradient pins them as it pins the Gaussian and negative-binomial ones."""
    eta = np.linspace(-2, 2, 9)
    h = 1e-6
    cases = [("gaussian", "identity", np.linspace(-1, 1, 9), 4.0, None,
              lambda e, y: ss.norm.logpdf(y, e, 0.5)),
             ("poisson", "log", np.arange(9.0), 2.5, None, lambda e, y: ss.poisson.logpmf(y, 2.5 * np.exp(e))),
             ("binomial", "logit", np.arange(9.0) % 6, 6.0, None,
              lambda e, y: ss.binom.logpmf(y, 6, 1 / (1 + np.exp(-e)))),
             ("nbinomial", "log", np.arange(9.0), 1.5, 5.0,
              lambda e, y: ss.nbinom.logpmf(y, 5, 5 / (5 + 1.5 * np.exp(e))))]
    for fam, link, y, a, b, ref in cases:
        lp, g, hs = orc.eval_lik(fam, link, eta, y, a, b)
        assert np.max(np.abs(lp - ref(eta, y))) < 1e-10, fam
        lpp, gp, _ = orc.eval_lik(fam, link, eta + h, y, a, b)
        lpm, gm, _ = orc.eval_lik(fam, link, eta - h, y, a, b)
        assert np.max(np.abs(g - (lpp - lpm) / (2 * h))) < 1e-5, fam
        assert np.max(np.abs(hs - (gp - gm) / (2 * h))) < 1e-4 * max(1.0, np.max(np.abs(hs))), fam


# ------------------------------------------------------------------ Newton / mlik
def pin_gaussian_closed_form():
    # inference.rs:2844-2929: 2 factorisations and 1 solve; mode = linear solve; the
    # Laplace mlik of a Gaussian model is exact: log N(y; 0, A Q^-1 A^T + I/prec)
    rng = np.random.default_rng(0)
    n, m = 30, 45
    q = M.ar1_q(n, 0.6, 2.0)
    A = sp.csc_matrix(rng.random((m, n)) * (rng.random((m, n)) < 0.1))
    y = rng.standard_normal(m)
    prec = 3.0
    r = orc.find_mode(q, "gaussian", "identity", y, prec, A=A)
    assert r.n_factorizations == 2 and r.n_solves == 1
    x = np.linalg.solve(q.toarray() + prec * (A.T @ A).toarray(), prec * A.T @ y)
    assert np.max(np.abs(r.x - x)) < 1e-10 * np.max(np.abs(x))
    S = A.toarray() @ np.linalg.solve(q.toarray() + 1e-12 * np.eye(n), A.T.toarray()) + np.eye(m) / prec
    assert abs(r.mlik - ss.multivariate_normal(np.zeros(m), S).logpdf(y)) < 1e-8


def pin_stationarity():
    """At the mode the gradient of sum log p(y | A x) - x^T Q x / 2 vanishes:
    A^T dlogp/deta - Q x* = 0, for each non-Gaussian family, with and without A."""
    rng = np.random.default_rng(1)
    n = 40
    q = M.ar1_q(n, 0.8, 1.5)
    m = 70
    A = sp.csc_matrix(rng.random((m, n)) * (rng.random((m, n)) < 0.08))
    eta0 = 0.3 * rng.standard_normal(m)
    cases = [("poisson", "log", rng.poisson(2.0, m).astype(float), np.full(m, 1.3), None),
             ("binomial", "logit", rng.binomial(8, 0.4, m).astype(float), np.full(m, 8.0), None),
             ("nbinomial", "log", rng.poisson(3.0, m).astype(float), np.full(m, 0.8), np.full(m, 4.0))]
    for fam, link, y, a, b in cases:
        for AA in (A, None):
            yy, aa, bb = (y, a, b) if AA is not None else (y[:n], a[:n], None if b is None else b[:n])
            r = orc.find_mode(q, fam, link, yy, aa, pb=bb, A=AA, tol=1e-12)
            assert r.n_factorizations > 2
            eta = AA @ r.x if AA is not None else r.x
            _, g, _ = orc.eval_lik(fam, link, eta, yy, aa, bb)
            grad = (AA.T @ g if AA is not None else g) - q @ r.x
            assert np.max(np.abs(grad)) < 1e-8 * max(1.0, np.max(np.abs(q @ r.x))), (fam, np.max(np.abs(grad)))


def pin_constrained():
    """Gaussian model with a sum-to-zero constraint through the reference's kappa
    augmentation and projection (constraints.rs:275-353), by dense brute force:
    x_u = (Q + kappa a a^T + prec I)^-1 prec y, x* = x_u - a (a^T a)^-1 a^T x_u, and
    mlik(x*) = log N(y; 0, (Q + kappa a a^T)^-1 + I/prec) - (x* - x_u)^T Q_post (x* - x_u)/2
    (the log-density is quadratic with its maximum at x_u). As kappa grows x* tends to
    the constrained conditional mean, which it matches to O(1/kappa). Two priors: an
    intrinsic RW1 (a spans its null space) and a proper AR1, for which a is not an
    eigenvector, so that kappa itself enters the mode."""
    n = 8
    rw1 = np.diag(np.r_[1.0, np.full(n - 2, 2.0), 1.0]) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    prec = 3.0
    ca = np.ones((1, n)) / math.sqrt(n)
    y = np.random.default_rng(2).standard_normal(n) + 0.4
    for Q in (2.0 * rw1, M.ar1_q(n, 0.6, 2.0).toarray()):
        r = orc.find_mode(sp.csc_matrix(Q), "gaussian", "identity", y, prec, constraints=(ca, [0.0]))
        assert abs(float((ca @ r.x)[0])) < 1e-12
        Qa = Q + KAPPA * ca.T @ ca
        Qp = Qa + prec * np.eye(n)
        xu = np.linalg.solve(Qp, prec * y)
        xs = xu - ca.T @ np.linalg.solve(ca @ ca.T, ca @ xu)
        assert np.max(np.abs(r.x - xs)) < 1e-10 * np.max(np.abs(xs))
        S = np.linalg.inv(Qa) + np.eye(n) / prec
        d = xs - xu
        exact = ss.multivariate_normal(np.zeros(n), S).logpdf(y) - 0.5 * d @ Qp @ d
        assert abs(r.mlik - exact) < 1e-6, (r.mlik, exact)  # kappa-conditioned dense inverse
        Sig = np.linalg.inv(Q + prec * np.eye(n))
        x0 = Sig @ (prec * y)
        xc = x0 - Sig @ ca.T @ np.linalg.solve(ca @ Sig @ ca.T, ca @ x0)
        assert np.max(np.abs(r.x - xc)) < 1e-5 * np.max(np.abs(xc))


# ------------------------------------------------------------------ integration
def pin_ccd():
    """integration.rs:1-62. Point counts (integration.rs:328-347); the weights are the
    ones for which the CCD mixture of a standard normal density reproduces its
    moments: sum W z = 0 and sum W z z^T = I with W_k = w_k exp(-|z_k|^2/2) normalised;
    m = 5 uses the resolution-V half fraction with defining relation I = +ABCDE."""
    for m, npts in ((1, 3), (2, 9), (3, 15), (4, 25), (5, 27)):
        for f0 in (1.0, 1.7):
            z, w = orc.ccd_design(m, f0)
            assert len(z) == npts and abs(w.sum() - 1.0) < 1e-12
            W = w * np.exp(-0.5 * np.sum(z * z, axis=1))
            W /= W.sum()
            assert np.max(np.abs(W @ z)) < 1e-13
            assert np.max(np.abs((z * W[:, None]).T @ z - np.eye(m))) < 1e-12
            radii = np.linalg.norm(z[1:], axis=1)
            assert np.allclose(radii, f0, rtol=0, atol=1e-14) and np.all(z[0] == 0)
    z, _ = orc.ccd_design(5, 1.0)
    fact = z[11:]
    assert len(fact) == 16
    assert np.all(np.prod(np.sign(fact), axis=1) == 1.0)
    assert len({tuple(np.sign(p)) for p in fact}) == 16


def pin_grid():
    """grid_design (integration.rs:64-156) on quadratic log-densities, where the point
    set is a closed-form lattice set: m = 1, g = -z^2/2: k step for |k| <= sqrt(2 thr)/step
    (centre, then the positive side outwards, then the negative side); m = 2,
    g = -(z1^2 + 4 z2^2)/2: the lattice points inside the ellipse, row by row."""
    thr = 4.0
    for step in (1.0, 0.7):
        pts, w = orc.grid_design(1, step, thr, lambda z: -0.5 * z @ z)
        K = int(math.floor(math.sqrt(2 * thr) / step + 1e-12))
        expect = [0.0] + [k * step for k in range(1, K + 1)] + [-k * step for k in range(1, K + 1)]
        assert len(pts) == len(expect) and np.allclose(pts[:, 0], expect, atol=1e-14) and np.all(w == 1.0)
    step = 0.8
    pts, _ = orc.grid_design(2, step, thr, lambda z: -0.5 * (z[0] ** 2 + 4 * z[1] ** 2))
    K1 = int(math.floor(math.sqrt(2 * thr) / step))
    K2 = int(math.floor(math.sqrt(2 * thr / 4) / step))
    expect = [(i * step, j * step) for i in range(-K1, K1 + 1) for j in range(-K2, K2 + 1)
              if 0.5 * ((i * step) ** 2 + 4 * (j * step) ** 2) <= thr]
    assert len(pts) == len(expect) and np.allclose(pts, np.array(expect), atol=1e-14)


def pin_invert():
    rng = np.random.default_rng(4)
    for a in (M.random_spd(5, 0.6, 1).toarray(), np.array([[0.0, 2.0, 1.0], [2.0, 1.0, 0.5], [1.0, 0.5, 3.0]])):
        inv = orc.invert_symmetric(a)
        assert np.max(np.abs(a @ inv - np.eye(len(a)))) < 1e-12
    with pytest.raises(orc.OracleError):
        orc.invert_symmetric(np.zeros((2, 2)))
    del rng


def pin_jacobi():
    rng = np.random.default_rng(6)
    for m in (2, 3, 5):
        g = rng.standard_normal((m, m))
        a = g @ g.T + np.eye(m)
        lam, V = orc.jacobi_eigen(a)
        assert np.max(np.abs(V.T @ V - np.eye(m))) < 1e-12
        assert np.max(np.abs(V @ np.diag(lam) @ V.T - a)) < 1e-11 * np.max(np.abs(a))
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(a), rtol=1e-12)


# ------------------------------------------------------------------ CPO / PIT
def pin_cdfs():
    """The CDFs of observation_cdf (model_selection.rs:449-503): the reference's known
    answers (model_selection.rs:768-791), scipy's Poisson and binomial CDFs, and for the
    normal formula the identity it follows: the rational term approximates
    erfc(|x|) exp(x^2) to 1.5e-7 (Abramowitz-Stegun 7.1.26), so with exp(-x^2 / 2) the
    formula is 1/2 (1 + sign(x) (1 - erfc(|x|) exp(x^2 / 2))) within 1.5e-7 exp(x^2 / 2)."""
    assert abs(orc.normal_cdf(0.0) - 0.5) < 1e-6 and orc.normal_cdf(3.0) > 0.998 and orc.normal_cdf(-3.0) < 0.002
    for x in np.linspace(-3.0, 3.0, 25):
        ident = 0.5 * (1.0 + math.copysign(1.0, x) * (1.0 - math.erfc(abs(x)) * math.exp(x * x / 2.0)))
        assert abs(orc.normal_cdf(x) - ident) <= 1.5e-7 * math.exp(x * x / 2.0) + 1e-15, x
    assert abs(orc.poisson_cdf(1.0, 0) - math.exp(-1.0)) < 1e-10 and orc.poisson_cdf(1.0, 10) > 0.9999
    for lam, k in ((0.3, 0), (2.5, 3), (7.0, 12), (40.0, 31)):
        assert abs(orc.poisson_cdf(lam, k) - ss.poisson.cdf(k, lam)) < 1e-12
    assert abs(orc.binomial_cdf(10, 0.5, 10) - 1.0) < 1e-10
    assert abs(orc.binomial_cdf(10, 0.5, 4) - 0.376953125) < 1e-8
    for n, p, k in ((10, 0.3, 2), (25, 0.8, 17), (7, 0.05, 0), (40, 0.6, 30)):
        assert abs(orc.binomial_cdf(n, p, k) - ss.binom.cdf(k, n, p)) < 1e-12


def pin_cpo_failure():
    """check_cpo_failure_for_obs: the reference's four cases (model_selection.rs:628-673),
    a non-positive variance and a leave-one-out precision <= 0."""
    assert not orc.cpo_failure("gaussian", "identity", 1.0, 2.0, 1.0, 0.2)
    assert orc.cpo_failure("gaussian", "identity", 1.0, 2.0, 1.0, 0.6)
    assert orc.cpo_failure("gaussian", "identity", 0.0, 5.0, 10.0, 0.1)
    assert orc.cpo_failure("poisson", "log", 1.0, 10.0, 0.0, 0.05)
    assert orc.cpo_failure("poisson", "log", 3.0, 1.0, 1.0, 0.0)
    assert orc.cpo_failure("poisson", "log", 3.0, 1.0, 1.0, 1.0 / np.e + 1e-3)  # 1/var < -hess = e
    assert not orc.cpo_failure("poisson", "log", 3.0, 1.0, 1.0, 0.1)


def pin_cpo():
    """compute_cpo_pit: CPO_i is the weighted harmonic mean over the nodes of the
    likelihood at the node predictors (scipy.stats.hmean); equal nodes give the
    likelihood itself; a single node fails (fewer than two terms); PIT of equal Gaussian
    nodes is the normal CDF of the standardised residual over 1/precision + variance."""
    rng = np.random.default_rng(7)
    n, K = 6, 4
    y = rng.poisson(3.0, size=n).astype(float)
    eta = np.log(3.0) + 0.05 * rng.standard_normal((K, n))
    var = np.full((K, n), 0.04)
    w = rng.uniform(0.5, 1.5, size=K)
    w /= w.sum()
    cpo, pit, nf = orc.cpo_pit("poisson", "log", y, eta, var, w, 1.0)
    assert nf == 0 and np.all(np.isfinite(cpo)) and np.all(np.isfinite(pit))
    lik = ss.poisson.pmf(y[None, :], np.exp(eta))
    assert np.allclose(cpo, [ss.hmean(lik[:, i], weights=w) for i in range(n)], rtol=1e-12)
    assert np.allclose(pit, np.clip(w @ ss.poisson.cdf(y[None, :], np.exp(eta)), 0, 1), rtol=1e-12)
    same = np.repeat(eta[:1], K, axis=0)
    cpo1, _, _ = orc.cpo_pit("poisson", "log", y, same, var, w, 1.0)
    assert np.allclose(cpo1, ss.poisson.pmf(y, np.exp(eta[0])), rtol=1e-12)
    c2, p2, nf2 = orc.cpo_pit("poisson", "log", y, eta[:1], var[:1], [1.0], 1.0)
    assert nf2 == n and np.all(np.isnan(c2)) and np.all(np.isnan(p2))
    yg = np.array([1.0, 1.1, 1.2])
    eg = np.repeat(np.array([[0.9, 1.15, 1.0]]), 3, axis=0)
    cg, pg, nfg = orc.cpo_pit("gaussian", "identity", yg, eg, np.full((3, 3), 0.2), [0.2, 0.5, 0.3], 2.0)
    assert nfg == 0
    assert np.allclose(pg, [orc.normal_cdf((yi - e) / math.sqrt(0.5 + 0.2)) for yi, e in zip(yg, eg[0])],
                       rtol=1e-14)
    # the negative binomial has no CDF: CPO only
    cn, pn, _ = orc.cpo_pit("nbinomial", "log", y, eta, var, w, 1.0, b=5.0)
    assert np.all(np.isfinite(cn)) and np.all(np.isnan(pn))


PINS = {k[4:]: v for k, v in dict(globals()).items() if k.startswith("pin_")}

# every exported oracle function with arithmetic, the mistake planted in it, and
# the pins that must catch it (orc_free only releases memory)
PLANT_PINS = {
    "lower_bw": ["bandwidths"], "bw_fwd": ["bandwidths"], "time_major": ["time_major"], "rcm": ["rcm"],
    "choose": ["chooser"], "min_degree": ["min_degree"], "ldlt": ["factor_brute"],
    "logdet": ["logdet_indefinite"], "solve": ["factor_brute"], "takahashi": ["factor_brute", "takahashi_known"],
    "selinv_lookup": ["selinv_entries"], "dense_ldlt": ["dense_ldlt"], "lgamma": ["lgamma"],
    "lik_gauss": ["likelihoods"], "lik_poisson": ["likelihoods"], "lik_binomial": ["likelihoods"],
    "lik_nbinomial": ["likelihoods"], "newton_rhs": ["stationarity"], "project": ["constrained"],
    "mlik": ["gaussian_closed_form"], "kappa": ["constrained"], "ccd5": ["ccd"], "ccd_weight": ["ccd"],
    "grid": ["grid"], "jacobi": ["jacobi"], "invert": ["invert"], "nnz_l": ["factor_brute"],
    "perm": ["factor_brute"], "d": ["factor_brute"], "cpo_fail": ["cpo_failure"], "cpo": ["cpo"],
    "normal_cdf": ["cdfs"], "poisson_cdf": ["cdfs", "cpo"], "binomial_cdf": ["cdfs"],
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


def test_every_plant_is_listed():
    assert sorted(PLANT_PINS) == sorted(orc.PLANTS[1:])


@pytest.mark.parametrize("plant", sorted(PLANT_PINS))
def test_planted_mistake_is_caught(plant):
    for pin in PLANT_PINS[plant]:
        with orc.planted(plant):
            with pytest.raises((AssertionError, orc.OracleError, np.linalg.LinAlgError)):
                PINS[pin]()
