"""Batched Laplace evaluation on the GPU vs the oracle's Newton loop at fixed theta.

Bars: Newton mode 1e-10 relative (max norm), marginal log-likelihood 1e-8
absolute, marginal variances 1e-9 relative; the adjoint gradient against central
finite differences (step 1e-4) of the oracle's marginal likelihood at 1e-5
relative (the FD truncation error dominates).
"""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MLIK_ATOL = 1e-8
MODE_RTOL = 1e-10
VAR_RTOL = 1e-9


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def spde_setup(nx, nobs, seed, family):
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(nx, nx)
    rng = np.random.default_rng(seed)
    A = W.lattice_projector(nx, nx, rng.uniform(size=(nobs, 2)))
    c, _ = W.spde_coefs(-0.5, 1.2)
    x = W.sample_gmrf(g.csc(c) + sp.identity(g.n) * 1e-9, rng)
    eta = A @ x
    if family == "gaussian":
        y = eta + 0.3 * rng.standard_normal(nobs)
        pa = np.ones(nobs)
    elif family == "poisson":
        y = rng.poisson(np.exp(0.5 + eta)).astype(float)
        pa = np.ones(nobs)
    elif family == "binomial":
        y = rng.binomial(10, 1 / (1 + np.exp(-eta))).astype(float)
        pa = np.full(nobs, 10.0)
    else:
        y = rng.negative_binomial(5, 5 / (5 + np.exp(0.5 + eta))).astype(float)
        pa = np.ones(nobs)
    return g, A, y, pa


def spde_params(theta_batch, gaussian):
    """theta = (log prec, log tau, log kappa) for Gaussian, else (log tau, log kappa)."""
    from inla_gpu import workloads as W

    coefs, dco, lp, dl = [], [], [], []
    for th in np.atleast_2d(theta_batch):
        off = 1 if gaussian else 0
        c, dc = W.spde_coefs(th[off], th[off + 1])
        d = np.zeros((len(th), 3))
        d[off:off + 2] = dc
        coefs.append(c)
        dco.append(d)
        lp.append(np.exp(th[0]) if gaussian else 1.0)
        e = np.zeros(len(th))
        if gaussian:
            e[0] = 1.0
        dl.append(e)
    return np.array(coefs), np.array(dco), np.array(lp), np.array(dl)


def oracle_mlik(g, A, y, pa, family, th, gaussian, pb=None, constraints=None, keep_post=False):
    c, _, lp, _ = spde_params(th, gaussian)
    q = g.csc(c[0])
    return orc.find_mode(q, family, {"gaussian": "identity", "poisson": "log", "binomial": "logit",
                                     "nbinomial": "log"}[family], y, pa * (lp[0] if gaussian else 1.0),
                         pb=pb, A=A, constraints=constraints, keep_post=keep_post)


@pytest.mark.parametrize("family", ["gaussian", "poisson", "binomial", "nbinomial"])
def test_laplace_matches_oracle(family, record_error):
    from inla_gpu.laplace import LaplaceModel

    gaussian = family == "gaussian"
    g, A, y, pa = spde_setup(18, 700, 7, family)
    pb = np.full(len(y), 5.0) if family == "nbinomial" else None
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, family=family, A=A, obs_a=pa,
                         obs_b=pb, capacity=4)
    thetas = np.array([[2.0, -0.5, 1.2], [2.5, -0.3, 1.0], [1.5, -0.8, 1.4]]) if gaussian else \
        np.array([[-0.5, 1.2], [-0.3, 1.0], [-0.8, 1.4]])
    coefs, dco, lp, dl = spde_params(thetas, gaussian)
    r = model.eval(coefs, lpar=lp, dcoefs=dco, dlog_lpar=dl)
    assert (r["status"] == 0).all(), r["status"]
    for b, th in enumerate(thetas):
        ref = oracle_mlik(g, A, y, pa, family, th, gaussian, pb=pb, keep_post=True)
        post = ref.post_factor()
        e_ml = record_error("mlik", abs(r["mlik"][b].item() - ref.mlik), MLIK_ATOL, "abs")
        e_x = record_error("mode", _rel(r["x"][b].cpu().numpy(), ref.x), MODE_RTOL)
        e_v = record_error("variances", _rel(r["var"][b].cpu().numpy(), post.takahashi_diag()), VAR_RTOL)
        assert e_ml < MLIK_ATOL, (b, r["mlik"][b].item(), ref.mlik)
        assert e_x < MODE_RTOL
        assert e_v < VAR_RTOL
        if not gaussian:
            assert r["n_iter"][b] >= 2


GRAD_RTOL = 1e-5  # central differences, step 1e-4: the truncation error dominates


def _central_fd(f, th, h=1e-4):
    fd = np.zeros_like(th)
    for l in range(len(th)):
        e = np.zeros_like(th)
        e[l] = h
        fd[l] = (f(th + e) - f(th - e)) / (2 * h)
    return fd


def _grad_err(grad, fd):
    return float(np.max(np.abs(grad - fd) / np.maximum(np.abs(fd), 1.0)))


@pytest.mark.parametrize("family", ["gaussian", "poisson", "binomial", "nbinomial"])
def test_gradient_matches_central_fd(family, record_error):
    from inla_gpu.laplace import LaplaceModel

    gaussian = family == "gaussian"
    g, A, y, pa = spde_setup(14, 400, 3, family)
    pb = np.full(len(y), 5.0) if family == "nbinomial" else None
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, family=family, A=A, obs_a=pa,
                         obs_b=pb, capacity=2)
    th = np.array([2.0, -0.5, 1.2]) if gaussian else np.array([-0.5, 1.2])
    coefs, dco, lp, dl = spde_params(th, gaussian)
    r = model.eval(coefs, lpar=lp, dcoefs=dco, dlog_lpar=dl, tol=1e-12)
    grad = r["grad"][0].cpu().numpy()
    fd = _central_fd(lambda t: oracle_mlik(g, A, y, pa, family, t, gaussian, pb=pb).mlik, th)
    err = record_error("gradient vs FD", _grad_err(grad, fd), GRAD_RTOL)
    assert err < GRAD_RTOL, (grad, fd)


def test_gaussian_log_link_gradient_matches_central_fd(record_error):
    """Gaussian likelihood with a log link: the Newton weights depend on eta, so the
    precision's gradient includes the mode's sensitivity to it (lambda^T A^T g)."""
    from inla_gpu.laplace import LaplaceModel

    g, A, _, _ = spde_setup(12, 300, 8, "gaussian")
    rng = np.random.default_rng(8)
    eta = A @ (0.3 * rng.standard_normal(g.n))
    y = np.exp(eta) + 0.2 * rng.standard_normal(len(eta))
    pa = np.ones(len(y))
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, family="gaussian", link="log", A=A,
                         obs_a=pa, capacity=1)
    th = np.array([1.5, -0.5, 1.2])
    coefs, dco, lp, dl = spde_params(th, True)
    r = model.eval(coefs, lpar=lp, dcoefs=dco, dlog_lpar=dl, tol=1e-12)
    assert r["status"][0] == 0

    def ml(t):
        c, _, lpt, _ = spde_params(t, True)
        return orc.find_mode(g.csc(c[0]), "gaussian", "log", y, pa * lpt[0], A=A, tol=1e-12).mlik

    ref = ml(th)
    e_ml = record_error("mlik (log link)", abs(r["mlik"][0].item() - ref), MLIK_ATOL, "abs")
    assert e_ml < MLIK_ATOL
    fd = _central_fd(ml, th)
    err = record_error("gradient vs FD (log link)", _grad_err(r["grad"][0].cpu().numpy(), fd), GRAD_RTOL)
    assert err < GRAD_RTOL, (r["grad"][0].cpu().numpy(), fd)


def test_log_marginal_autograd_matches_central_fd(record_error):
    """LogMarginal (torch.autograd.Function): backward of a weighted sum of mlik over
    a batch of theta-points against central differences of the oracle's mlik."""
    import torch

    from inla_gpu.laplace import LaplaceModel, LogMarginal

    g, A, y, pa = spde_setup(12, 300, 5, "poisson")
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, family="poisson", A=A, obs_a=pa,
                         capacity=2)

    def param_fn(th):
        c, dc, lp, dl = spde_params(th, False)
        return c, dc, lp, dl

    theta = torch.tensor([[-0.5, 1.2], [-0.3, 1.0]], dtype=torch.float64, requires_grad=True)
    w = torch.tensor([1.0, -2.0], dtype=torch.float64)
    out = LogMarginal.apply(theta, model, param_fn)
    (out * w).sum().backward()
    got = theta.grad.numpy()
    for b in range(2):
        th = theta.detach().numpy()[b]
        ref = oracle_mlik(g, A, y, pa, "poisson", th, False).mlik
        e_ml = record_error("mlik (autograd forward)", abs(out[b].item() - ref), MLIK_ATOL, "abs")
        assert e_ml < MLIK_ATOL
        fd = w[b].item() * _central_fd(lambda t: oracle_mlik(g, A, y, pa, "poisson", t, False).mlik, th)
        err = record_error("autograd grad vs FD", _grad_err(got[b], fd), 1e-4)
        assert err < 1e-4, (got[b], fd)


def test_gaussian_mlik_closed_form(record_error):
    # Gaussian likelihood: the Laplace mlik is exact, log N(y; 0, A Q^-1 A^T + I/prec)
    from inla_gpu.laplace import LaplaceModel

    g, A, y, pa = spde_setup(8, 90, 11, "gaussian")
    th = np.array([1.0, -0.2, 1.1])
    coefs, _, lp, _ = spde_params(th, True)
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, A=A, obs_a=pa, capacity=1)
    r = model.eval(coefs, lpar=lp)
    Q = g.csc(coefs[0]).toarray() + 1e-12 * np.eye(g.n)
    Ad = A.toarray()
    S = Ad @ np.linalg.solve(Q, Ad.T) + np.eye(len(y)) / lp[0]
    sign, logdet = np.linalg.slogdet(S)
    exact = -0.5 * (len(y) * np.log(2 * np.pi) + logdet + y @ np.linalg.solve(S, y))
    # the dense closed form carries numpy's own rounding on an ill-conditioned
    # covariance (the prior is intrinsic up to the 1e-12 jitter)
    err = record_error("mlik vs closed form", abs(r["mlik"][0].item() - exact), 1e-7, "abs")
    assert err < 1e-7


def test_constrained_besag_matches_oracle(record_error):
    # Besag ICAR + sum-to-zero, kappa A_c^T A_c augmentation (constraints.rs:275-313)
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    nr = 7
    adj = W.rook_adjacency(nr, nr)
    R = W.besag_structure(adj)
    n = R.shape[0]
    kappa = 3.059023205018258e6
    ca = np.full((1, n), 1.0 / np.sqrt(n))
    Kc = sp.csc_matrix(kappa * ca.T @ ca)
    indptr, indices = W.union_pattern([R, Kc])
    comps = np.stack([W.values_on_pattern(R, indptr, indices), W.values_on_pattern(Kc, indptr, indices)])
    rng = np.random.default_rng(4)
    y = rng.poisson(np.exp(0.5 + 0.3 * rng.standard_normal(n))).astype(float)
    model = LaplaceModel(n, indptr, indices, comps, y, family="poisson", constraints=(ca, [0.0]),
                         capacity=2)
    taus = [1.0, 3.0]
    r = model.eval(np.array([[t, 1.0] for t in taus]))
    assert (r["status"] == 0).all()
    for b, tau in enumerate(taus):
        ref = orc.find_mode(R * tau, "poisson", "log", y, np.ones(n), constraints=(ca, [0.0]))
        e_ml = record_error("mlik (kappa-augmented)", abs(r["mlik"][b].item() - ref.mlik), MLIK_ATOL, "abs")
        e_x = record_error("mode (kappa-augmented)", _rel(r["x"][b].cpu().numpy(), ref.x), MODE_RTOL)
        assert e_ml < MLIK_ATOL, (r["mlik"][b].item(), ref.mlik)
        assert e_x < MODE_RTOL


def test_fit_batched_and_plain_agree():
    """End-to-end fit: batched BFGS on the adjoint gradient and the reference's
    one-point Nelder-Mead sequence reach the same mode and mixture; the batched
    path evaluates far fewer calls."""
    from inla_gpu import fit as F
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload(nx=20, n_obs=600, seed=5)
    g = wl.gmrf
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A,
                         capacity=19)
    theta0 = wl.theta_true + np.array([0.4, -0.3, 0.2])
    rb = F.fit_batched(F.Objective(model, wl.params, wl.log_prior), theta0, gtol=1e-4)
    rp = F.fit_plain(F.Objective(model, wl.params, wl.log_prior), theta0)
    print("modes", rb.mode, rp.mode, "calls", rb.n_evals, rp.n_evals)
    assert np.max(np.abs(rb.mode - rp.mode)) < 2e-2
    assert abs(rb.mlik - rp.mlik) < 1e-2
    assert _rel(rb.mean, rp.mean) < 2e-2
    assert rb.n_evals < rp.n_evals
    # the adjoint gradient vanishes at the batched mode
    ob = F.Objective(model, wl.params, wl.log_prior)
    gr = ob.values(rb.mode[None], grad=True)["grad"][0]
    assert np.max(np.abs(gr)) < 1e-3


def test_spde_poisson_with_intercept_matches_oracle(record_error):
    """The batched Poisson configuration (intercept + SPDE, dense intercept row in
    Q_post) at small size: mode, mlik and variances against the oracle."""
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload(nx=12, n_obs=300, seed=9)
    g = wl.gmrf
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0], [0.2, -0.1], [-0.15, 0.1]])
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         capacity=len(thetas))
    c, _, _, _ = wl.params(thetas)
    r = model.eval(c)
    assert np.all(r["status"] == 0)
    for b in range(len(thetas)):
        ref = orc.find_mode(g.csc(c[b]), "poisson", "log", wl.y, np.ones(len(wl.y)), A=wl.A, keep_post=True)
        e_ml = record_error("mlik", abs(r["mlik"][b].item() - ref.mlik), MLIK_ATOL, "abs")
        e_x = record_error("mode", _rel(r["x"][b].cpu().numpy(), ref.x), MODE_RTOL)
        e_v = record_error("variances", _rel(r["var"][b].cpu().numpy(), ref.post_factor().takahashi_diag()),
                           VAR_RTOL)
        assert e_ml < MLIK_ATOL
        assert e_x < MODE_RTOL
        assert e_v < VAR_RTOL


def test_spde_poisson_long_gather_rows_match_oracle(record_error):
    """More observations than the per-thread gather limit (512): the intercept's
    column of A (A^T g, A^T u) and its diagonal entry of A^T W A (every
    observation) are summed by one workgroup each. Mode, mlik and variances against
    the oracle; the adjoint gradient against central differences of the oracle's
    mlik (1e-5 relative, as the other gradient tests)."""
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload(nx=12, n_obs=2000, seed=21)
    g = wl.gmrf
    assert np.diff(wl.A.indptr).max() > 512  # the intercept column
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0], [0.15, -0.1]])
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         capacity=len(thetas))
    c, dc, _, _ = wl.params(thetas)
    r = model.eval(c, dcoefs=dc, tol=1e-12)
    assert np.all(r["status"] == 0)
    ones = np.ones(len(wl.y))
    for b in range(len(thetas)):
        ref = orc.find_mode(g.csc(c[b]), "poisson", "log", wl.y, ones, A=wl.A, tol=1e-12, keep_post=True)
        e_ml = record_error("mlik", abs(r["mlik"][b].item() - ref.mlik), MLIK_ATOL, "abs")
        e_x = record_error("mode", _rel(r["x"][b].cpu().numpy(), ref.x), MODE_RTOL)
        e_v = record_error("variances", _rel(r["var"][b].cpu().numpy(), ref.post_factor().takahashi_diag()),
                           VAR_RTOL)
        assert e_ml < MLIK_ATOL
        assert e_x < MODE_RTOL
        assert e_v < VAR_RTOL
    th = thetas[1]
    h = 1e-4
    fd = np.zeros(2)
    for l in range(2):
        vals = []
        for sgn in (1.0, -1.0):
            t = th.copy()
            t[l] += sgn * h
            cc, _, _, _ = wl.params(t[None])
            vals.append(orc.find_mode(g.csc(cc[0]), "poisson", "log", wl.y, ones, A=wl.A, tol=1e-12).mlik)
        fd[l] = (vals[0] - vals[1]) / (2 * h)
    grad = r["grad"][1].cpu().numpy()
    err = record_error("gradient vs FD", _grad_err(grad, fd), GRAD_RTOL)
    assert err < GRAD_RTOL, (grad, fd)
