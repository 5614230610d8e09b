"""Full-size SPDE Gaussian configuration (142 x 142 lattice, n = 20,164, 50,000
observations, CCD m = 3: 15 theta-points) through the batched Laplace path against
the oracle, point by point: marginal log-likelihood (1e-8 absolute), mode (1e-10
relative), marginal variances (1e-9 relative), and the adjoint gradient at the
design's centre against central differences of the oracle's mlik (step 1e-4,
1e-5 relative). The oracle's points run in a thread pool (its calls release the
interpreter lock)."""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MLIK_ATOL = 1e-8
MODE_RTOL = 1e-10
VAR_RTOL = 1e-9
GRAD_RTOL = 1e-5


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def test_spde_gaussian_full_size_ccd_batch_matches_oracle(record_error):
    from inla_gpu import ccd
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload()
    g = wl.gmrf
    assert g.n == 20164 and len(wl.y) == 50_000
    m = 3
    hh = max(1e-4, 0.05 * float(np.max(np.abs(wl.theta_true))))
    stencil = ccd.hessian_points(wl.theta_true, hh)
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A,
                         capacity=len(stencil))
    c, _, lp, _ = wl.params(stencil)
    r = model.eval(c, lpar=lp, want_x=False, want_var=False)
    lpost = r["mlik"].cpu().numpy() + np.array([wl.log_prior(t)[0] for t in stencil])
    lam, V = ccd.scaling_from_hessian(ccd.hessian_from_values(lpost, m, hh))
    z, _ = ccd.ccd_design(m, 1.0)
    thetas = np.array([ccd.theta_from_z(wl.theta_true, lam, V, zz) for zz in z])
    assert len(thetas) == 15
    coefs, dcoefs, lpar, dlog = wl.params(thetas)
    out = model.eval(coefs, lpar=lpar, dcoefs=dcoefs, dlog_lpar=dlog)
    assert (out["status"] == 0).all(), out["status"]
    mlik = out["mlik"].cpu().numpy()
    x = out["x"].cpu().numpy()
    var = out["var"].cpu().numpy()
    grad = out["grad"].cpu().numpy()

    def oracle_point(b):
        ref = orc.find_mode(g.csc(coefs[b]), "gaussian", "identity", wl.y, np.full(len(wl.y), lpar[b]),
                            A=wl.A, keep_post=True)
        return ref.mlik, ref.x, ref.post_factor().takahashi_diag()

    def oracle_mlik(th):
        cc, _, lpp, _ = wl.params(np.asarray(th)[None])
        return orc.find_mode(g.csc(cc[0]), "gaussian", "identity", wl.y, np.full(len(wl.y), lpp[0]),
                             A=wl.A).mlik

    h = 1e-4
    centre = thetas[0]
    fd_points = []
    for l in range(m):
        for s in (h, -h):
            t = centre.copy()
            t[l] += s
            fd_points.append(t)
    with ThreadPoolExecutor(_workers()) as ex:
        refs = list(ex.map(oracle_point, range(len(thetas))))
        fdv = list(ex.map(oracle_mlik, fd_points))
    e_ml = max(abs(mlik[b] - refs[b][0]) for b in range(15))
    e_x = max(_rel(x[b], refs[b][1]) for b in range(15))
    e_v = max(_rel(var[b], refs[b][2]) for b in range(15))
    fd = np.array([(fdv[2 * l] - fdv[2 * l + 1]) / (2 * h) for l in range(m)])
    e_g = float(np.max(np.abs(grad[0] - fd) / np.maximum(np.abs(fd), 1.0)))
    record_error("mlik (15 points)", e_ml, MLIK_ATOL, "abs")
    record_error("mode (15 points)", e_x, MODE_RTOL)
    record_error("variances (15 points)", e_v, VAR_RTOL)
    record_error("gradient vs FD (centre)", e_g, GRAD_RTOL)
    assert e_ml < MLIK_ATOL
    assert e_x < MODE_RTOL
    assert e_v < VAR_RTOL
    assert e_g < GRAD_RTOL, (grad[0], fd)


def test_spde_poisson_full_size_ccd_batch_matches_oracle(record_error):
    """The batched Poisson configuration at full size (the 142 x 142 SPDE mesh plus an
    intercept, n = 20,165, 50,000 counts y ~ Poisson(exp(1 + u)), CCD m = 2 around the
    generating theta with unit scaling: 9 points, each its own Newton loop) against the
    oracle point by point, and the adjoint gradient at the centre against central
    differences of the oracle's mlik."""
    from inla_gpu import ccd
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload()
    g = wl.gmrf
    assert g.n == 20165 and len(wl.y) == 50_000
    z, _ = ccd.ccd_design(2, 1.0)
    thetas = wl.theta_true[None] + 0.3 * z
    coefs, dcoefs, _, _ = wl.params(thetas)
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         capacity=len(thetas))
    out = model.eval(coefs, dcoefs=dcoefs)
    assert (out["status"] == 0).all(), out["status"]
    mlik, x, var, grad = (out[k].cpu().numpy() for k in ("mlik", "x", "var", "grad"))

    def oracle_point(b):
        ref = orc.find_mode(g.csc(coefs[b]), "poisson", "log", wl.y, 1.0, A=wl.A, keep_post=True)
        return ref.mlik, ref.x, ref.post_factor().takahashi_diag()

    def oracle_mlik(th):
        cc, _, _, _ = wl.params(np.asarray(th)[None])
        return orc.find_mode(g.csc(cc[0]), "poisson", "log", wl.y, 1.0, A=wl.A).mlik

    h = 1e-4
    fd_points = []
    for l in range(2):
        for s in (h, -h):
            t = thetas[0].copy()
            t[l] += s
            fd_points.append(t)
    with ThreadPoolExecutor(_workers()) as ex:
        refs = list(ex.map(oracle_point, range(len(thetas))))
        fdv = list(ex.map(oracle_mlik, fd_points))
    for b, (ml, xo, vo) in enumerate(refs):
        e_ml = record_error("mlik (Poisson, 9 points)", abs(mlik[b] - ml), MLIK_ATOL, "abs")
        e_x = record_error("mode (Poisson, 9 points)", _rel(x[b], xo), MODE_RTOL)
        e_v = record_error("variances (Poisson, 9 points)", _rel(var[b], vo), VAR_RTOL)
        assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (b, e_ml, e_x, e_v)
    fd = np.array([(fdv[2 * l] - fdv[2 * l + 1]) / (2 * h) for l in range(2)])
    e_g = record_error("gradient vs FD (Poisson, centre)", float(np.max(np.abs(grad[0] - fd) / np.maximum(np.abs(fd), 1.0))),
                       GRAD_RTOL)
    assert e_g < GRAD_RTOL, (grad[0], fd)
