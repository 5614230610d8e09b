"""Full-size areal Poisson configuration (30 x 30 rook Besag ICAR with scale_model,
intercept and covariate, sum-to-zero on u, n = 902; theta = log tau with the
loggamma(1, 5e-5) prior; 1-D grid integration) against the oracle:

* the Laplace evaluation in the reference's form (kappa a a^T as a component: a
  dense 902 x 902 front) and in the low-rank form, at three theta: mlik (1e-8
  absolute), mode (1e-10 relative), marginal variances (1e-9 relative), and the
  adjoint gradient against central differences of the oracle's mlik;
* the grid integration: the grid points from the oracle's log posterior equal the
  GPU fit's, and the mixture means and variances agree.

The oracle's evaluations run in a thread pool (its calls release the interpreter lock).
"""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MLIK_ATOL = 1e-8
MODE_RTOL = 1e-10
VAR_RTOL = 1e-9


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def _pool():
    return ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0)))))


@pytest.fixture(scope="module")
def c1():
    from inla_gpu import workloads as W

    wl = W.besag_poisson_workload()
    assert wl.n == 902 and len(wl.gmrf.indices) == 902 * 902
    return wl


def _oracle(wl, th, keep_post=False, tol=1e-5):
    ca, ce = wl.constraints
    return orc.find_mode(wl.prior_without_kappa(th), "poisson", "log", wl.y, np.ones(len(wl.y)), A=wl.A,
                         constraints=(ca, ce), keep_post=keep_post, tol=tol)


def test_c1_laplace_matches_oracle_both_constraint_forms(c1, record_error):
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = c1
    ca, ce = wl.constraints
    thetas = np.array([[0.0], [0.8], [-0.6]])
    g, g2 = wl.gmrf, wl.gmrf_lowrank
    dense = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         constraints=(ca, ce), capacity=3)
    lowr = LaplaceModel(g2.n, g2.indptr, g2.indices, g2.components, wl.y, family="poisson", A=wl.A,
                        constraints=(ca, ce), constraint_kappa=W.KAPPA_HARD, capacity=3)
    c, dc, _, _ = wl.params(thetas)
    rd = dense.eval(c, dcoefs=dc)
    c2, dc2, _, _ = wl.params_lowrank(thetas)
    rl = lowr.eval(c2, dcoefs=dc2)
    assert (rd["status"] == 0).all() and (rl["status"] == 0).all()

    def ref(th):
        r = _oracle(wl, th, keep_post=True)
        return r.mlik, r.x, r.post_factor().takahashi_diag()

    h = 1e-4
    fd_pts = [thetas[0, 0] + h, thetas[0, 0] - h]
    with _pool() as ex:
        refs = list(ex.map(ref, thetas[:, 0]))
        fdv = list(ex.map(lambda t: _oracle(wl, t).mlik, fd_pts))
    for name, r in (("dense", rd), ("low-rank", rl)):
        for b in range(len(thetas)):
            e_ml = record_error(f"mlik ({name})", abs(r["mlik"][b].item() - refs[b][0]), MLIK_ATOL, "abs")
            e_x = record_error(f"mode ({name})", _rel(r["x"][b].cpu().numpy(), refs[b][1]), MODE_RTOL)
            e_v = record_error(f"variances ({name})", _rel(r["var"][b].cpu().numpy(), refs[b][2]), VAR_RTOL)
            assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (name, b, e_ml, e_x, e_v)
            assert abs(float((ca @ r["x"][b].cpu().numpy())[0])) < 1e-12
        fd = (fdv[0] - fdv[1]) / (2 * h)
        gr = r["grad"][0].item()
        e_g = record_error(f"gradient vs FD ({name})", abs(gr - fd) / max(abs(fd), 1.0), 1e-5)
        assert e_g < 1e-5, (name, gr, fd)


def test_c1_grid_integration_matches_oracle(c1, record_error):
    from inla_gpu import ccd
    from inla_gpu import fit as F
    from inla_gpu.laplace import LaplaceModel

    wl = c1
    ca, ce = wl.constraints
    g = wl.gmrf
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         constraints=(ca, ce), capacity=8)
    rb = F.fit_batched(F.Objective(model, wl.params, wl.log_prior), wl.theta_true + 0.4, strategy="grid")
    lam, V = ccd.scaling_from_hessian(rb.hessian)
    sd = float(np.sqrt(abs(lam[0])) * V[0, 0])

    def log_post(th):
        return _oracle(wl, th).mlik + wl.log_prior([th])[0]

    # the oracle's log posterior on the z-lattice the walk can reach, in parallel
    zs = np.arange(-6, 7) * 1.0
    with _pool() as ex:
        gz = dict(zip(zs.tolist(), ex.map(lambda z: log_post(rb.mode[0] + sd * z), zs)))
    zo, wo = orc.grid_design(1, 1.0, F.GRID_THRESHOLD, lambda z: gz[float(np.round(z[0], 12))])
    z_gpu = (rb.thetas[:, 0] - rb.mode[0]) / sd
    assert len(zo) == len(z_gpu) and np.allclose(zo[:, 0], z_gpu, atol=1e-12), (zo[:, 0], z_gpu)
    # the mixture from the oracle's nodes
    with _pool() as ex:
        nodes = list(ex.map(lambda th: _oracle(wl, th, keep_post=True), rb.thetas[:, 0]))
        dvar = list(ex.map(lambda r: r.post_factor().takahashi_diag(), nodes))
    lp = np.array([r.mlik + wl.log_prior([t])[0] for r, t in zip(nodes, rb.thetas[:, 0])])
    xs = np.array([r.x for r in nodes])
    wts, mean, var, mlik = ccd.aggregate(wo, lp, xs, np.array(dvar))
    e_lp = record_error("grid log posterior", float(np.max(np.abs(rb.log_post - lp))), MLIK_ATOL, "abs")
    # a log-posterior error eps moves a normalised weight by at most about 2 eps
    e_w = record_error("grid weights", float(np.max(np.abs(rb.weights - wts))), 2 * MLIK_ATOL, "abs")
    e_m = record_error("mixture mean", _rel(rb.mean, mean), 1e-9)
    e_v = record_error("mixture variance", _rel(rb.var, var), 1e-9)
    e_ml = record_error("mixture mlik", abs(rb.mlik - mlik), MLIK_ATOL, "abs")
    assert e_lp < MLIK_ATOL and e_w < 2 * MLIK_ATOL and e_m < 1e-9 and e_v < 1e-9 and e_ml < MLIK_ATOL
