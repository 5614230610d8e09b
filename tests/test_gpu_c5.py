"""The large combined configuration's model at test size: SPDE x RW2 space-time model
with five hyperparameters (Q = Q_spde(tau1, kappa1) x (R_rw2 + eps I) +
Q_spde(tau2, kappa2) x I_T + tau3 I, seven affine components), negative binomial
(size 5) counts, against the oracle:

* the Laplace evaluation at three theta-points: mlik (1e-8 absolute), mode (1e-10
  relative), marginal variances (1e-9 relative); the lean form (prior factorised alone
  first, as the full size runs) equals the standard form; the adjoint gradient's five
  components against central differences (step 1e-4) of the oracle's mlik;
* the whole fit: batched BFGS on the adjoint gradient, the Hessian stencil and the
  m = 5 CCD design (27 nodes: centre, 10 axial points, the 16-point half fraction of
  crates/inla_math/src/integration.rs:31-44) in one batch; every node's log posterior
  against the oracle's, and the mixture's mean and variance against the mixture of the
  oracle's node modes and Takahashi diagonals.
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
def c5():
    from inla_gpu import workloads as W

    return W.spde_rw2_five_workload(nx=8, T=4, n_obs=400, seed=17)


def _model(wl, capacity, lean=False):
    from inla_gpu.laplace import LaplaceModel

    g = wl.gmrf
    return LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="nbinomial", A=wl.A,
                        obs_a=wl.obs_a, obs_b=wl.obs_b, capacity=capacity, lean=lean)


def _oracle(wl, th, keep_post=False):
    c, _, _, _ = wl.params(th[None])
    return orc.find_mode(wl.gmrf.csc(c[0]), "nbinomial", "log", wl.y, wl.obs_a, pb=wl.obs_b, A=wl.A,
                         keep_post=keep_post)


def test_c5_five_hyperparameters_laplace_matches_oracle(c5, record_error):
    wl = c5
    assert wl.gmrf.components.shape[0] == 7 and len(wl.theta_true) == 5
    thetas = wl.theta_true[None] + np.array([[0.0] * 5, [0.2, -0.15, 0.1, 0.1, -0.2], [-0.1, 0.1, -0.2, 0.0, 0.3]])
    c, dc, _, _ = wl.params(thetas)
    r = _model(wl, 3).eval(c, dcoefs=dc)
    assert np.all(r["status"] == 0)
    rl = _model(wl, 3, lean=True).eval(c, dcoefs=dc)
    assert np.all(rl["status"] == 0)
    for k in ("mlik", "x", "var", "grad"):
        a, b = rl[k].cpu().numpy(), r[k].cpu().numpy()
        record_error(f"lean vs standard {k}", _rel(a, b), 1e-12)
        assert _rel(a, b) < 1e-12, k

    def ref(th):
        o = _oracle(wl, th, keep_post=True)
        return o.mlik, o.x, o.post_factor().takahashi_diag()

    h = 1e-4
    fd_pts = []
    for l in range(5):
        for sgn in (h, -h):
            t = thetas[1].copy()
            t[l] += sgn
            fd_pts.append(t)
    with _pool() as ex:
        refs = list(ex.map(ref, thetas))
        fdv = list(ex.map(lambda t: _oracle(wl, t).mlik, fd_pts))
    for b, (ml, x, v) in enumerate(refs):
        e_ml = record_error("mlik", abs(r["mlik"][b].item() - ml), MLIK_ATOL, "abs")
        e_x = record_error("mode", _rel(r["x"][b].cpu().numpy(), x), MODE_RTOL)
        e_v = record_error("variances", _rel(r["var"][b].cpu().numpy(), v), VAR_RTOL)
        assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (b, e_ml, e_x, e_v)
    fd = np.array([(fdv[2 * l] - fdv[2 * l + 1]) / (2 * h) for l in range(5)])
    g = r["grad"][1].cpu().numpy()
    err = record_error("gradient vs FD (5 components)", _rel(g, fd), 1e-5)
    assert err < 1e-5, (g, fd)


def test_c5_five_hyperparameters_ccd_fit_matches_oracle(c5, record_error):
    from inla_gpu import ccd
    from inla_gpu.fit import Objective, fit_batched

    wl = c5
    model = _model(wl, 27)
    obj = Objective(model, wl.params, wl.log_prior)
    res = fit_batched(obj, wl.theta_true + 0.2)
    assert len(res.thetas) == 27
    z, _ = ccd.ccd_design(5, 1.0)
    frac = z[11:]
    assert len(frac) == 16 and np.allclose(np.prod(np.sign(frac), axis=1), 1.0)
    assert np.all(np.linalg.eigvalsh(res.hessian) < 0)

    def node(th):
        o = _oracle(wl, th, keep_post=True)
        return o.mlik + wl.log_prior(th)[0], o.x, o.post_factor().takahashi_diag()

    with _pool() as ex:
        refs = list(ex.map(node, res.thetas))
    lp = np.array([q[0] for q in refs])
    err = record_error("27 node log posteriors", float(np.max(np.abs(res.log_post - lp))), MLIK_ATOL, "abs")
    assert err < MLIK_ATOL
    _, w = ccd.ccd_design(5, 1.0)
    wts, mean, var, _ = ccd.aggregate(w, lp, np.array([q[1] for q in refs]), np.array([q[2] for q in refs]))
    e_w = record_error("mixture weights", float(np.max(np.abs(res.weights - wts))), 1e-8, "abs")
    e_m = record_error("mixture mean", _rel(res.mean, mean), 1e-9)
    e_v = record_error("mixture variance", _rel(res.var, var), 1e-8)
    assert e_w < 1e-8 and e_m < 1e-9 and e_v < 1e-8
