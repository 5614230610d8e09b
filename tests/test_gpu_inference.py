"""run_inla_inference (py-inla/src/lib.rs:427-657) on the GPU against the same driver
run over the oracle's node evaluations (find_mode + Takahashi per theta, the reference's
per-node work): an areal Poisson model given as build_prior / log_prior_density callbacks
and observation dicts, with a sum-to-zero constraint and the grid strategy (the C1 shape
at 8 x 8), and a Gaussian SPDE model with a free observation precision and the CCD
strategy. Both runs follow the same search (Nelder-Mead), so every output must agree to
the parity bars: mode 1e-6 (a Nelder-Mead tolerance of 1e-6 on values that agree to
1e-9), log posteriors and mlik 1e-8 absolute, means / variances / predictor moments 1e-8
relative, DIC / WAIC 1e-7 absolute, CPO / PIT 1e-8 relative with the same None pattern.
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KAPPA = 3.059023205018258e6


class OracleNodes:
    """The node-evaluator interface of inla_gpu.inference.GpuNodes on the oracle."""

    def __init__(self, gn):
        self.gn = gn
        self.calls = self.points = 0

    def lpar(self, th):
        return self.gn.lpar(th)

    def values(self, thetas, grad=False, want_x=False, want_var=False, want_eta=False):
        assert not grad
        gn = self.gn
        thetas = np.atleast_2d(thetas)
        out = {k: [] for k in ("g", "x", "var", "eta", "eta_var", "status")}
        A = gn.model.A
        cons = None
        if gn.kappa_aug is not None:
            cons = (self.ca, self.ce)
        for th in thetas:
            # (the oracle's find_mode adds the kappa augmentation itself, constraints.rs:275-313)
            q = sp.csc_matrix(gn.build_prior(th))
            lp, _ = gn.lpar(th[None])
            pa = gn.obs_a * (lp[0] if lp is not None else 1.0)
            r = orc.find_mode(q, gn.family, gn.link, gn.y, pa, pb=gn.obs_b, A=A, constraints=cons,
                              keep_post=want_var or want_eta)
            out["g"].append(r.mlik + float(gn.log_prior_density(th)))
            out["status"].append(0)
            x = r.x
            out["x"].append(x)
            if want_var or want_eta:
                var = r.post_factor().takahashi_diag()
                out["var"].append(var)
                out["eta"].append(A @ x if A is not None else x)
                out["eta_var"].append(A.multiply(A) @ var if A is not None else var)
        self.calls += 1
        self.points += len(thetas)
        return {k: np.array(v) for k, v in out.items() if v}


def _besag_model(nr=8, seed=3):
    from inla_gpu import workloads as W

    adj = W.rook_adjacency(nr, nr)
    R = W.besag_structure(adj)
    ns = nr * nr
    Qs = sp.csc_matrix(R * W.scale_factor_dense(R))
    rng = np.random.default_rng(seed)
    xc = rng.standard_normal(ns)
    y = rng.poisson(np.exp(0.5 + 0.3 * xc + 0.4 * rng.standard_normal(ns))).astype(float)
    n = ns + 2
    A = sp.csc_matrix((np.r_[np.ones(ns), xc, np.ones(ns)],
                       (np.r_[np.arange(ns), np.arange(ns), np.arange(ns)],
                        np.r_[np.zeros(ns, int), np.ones(ns, int), 2 + np.arange(ns)])), shape=(ns, n))

    def build_prior(theta):
        return sp.csc_matrix(sp.block_diag([sp.diags([1e-4, 1e-4]), math.exp(theta[0]) * Qs]))

    def log_prior(theta):
        return W.loggamma_prior(theta[0])[0]

    ca = np.zeros((1, n))
    ca[0, 2:] = 1.0 / math.sqrt(ns)
    obs = [{"family": "poisson", "y": v, "E": 1.0} for v in y]
    return build_prior, log_prior, obs, A, ca


def _compare(rg, ro, record_error, tag):
    assert np.max(np.abs(rg["mode"] - ro["mode"])) < 1e-6
    assert np.allclose(rg["theta_nodes"], ro["theta_nodes"], rtol=0, atol=1e-5)
    e_lp = record_error(f"front-end log posterior ({tag})", float(np.max(np.abs(rg["log_posts"] - ro["log_posts"]))),
                        1e-8, "abs")
    e_ml = record_error(f"front-end mlik ({tag})", abs(rg["marginal_log_lik"] - ro["marginal_log_lik"]), 1e-8, "abs")
    assert e_lp < 1e-8 and e_ml < 1e-8
    for k in ("latent_means", "latent_variances", "predictor_means", "predictor_variances", "node_weights"):
        e = float(np.max(np.abs(rg[k] - ro[k])) / np.max(np.abs(ro[k])))
        record_error(f"front-end {k} ({tag})", e, 1e-8)
        assert e < 1e-8, (k, e)
    for k in ("dic", "mean_deviance", "waic", "waic_lppd"):
        assert abs(rg[k] - ro[k]) < 1e-7, (k, rg[k], ro[k])
    assert rg["cpo_n_failures"] == ro["cpo_n_failures"]
    for k in ("cpo", "pit"):
        a = np.array([np.nan if v is None else v for v in rg[k]])
        b = np.array([np.nan if v is None else v for v in ro[k]])
        assert np.array_equal(np.isnan(a), np.isnan(b))
        ok = ~np.isnan(b)
        if ok.any():
            assert np.max(np.abs(a[ok] - b[ok]) / np.abs(b[ok])) < 1e-8
    for mg, mo in zip(rg["marginals_latent"] + rg["internal_marginals_hyperpar"],
                      ro["marginals_latent"] + ro["internal_marginals_hyperpar"]):
        assert np.allclose(mg.x, mo.x, rtol=1e-6, atol=1e-9) and np.allclose(mg.y, mo.y, rtol=1e-5, atol=1e-9)


def test_front_end_areal_poisson_grid_matches_oracle_nodes(record_error):
    from inla_gpu import inference as I

    build_prior, log_prior, obs, A, ca = _besag_model()
    fam, link, y, oa, ob, slot = I.parse_obs(obs)
    gn = I.GpuNodes(build_prior, log_prior, 1, fam, link, y, oa, ob, slot, A=A, constraints=(ca, np.zeros(1)),
                    capacity=8, theta0=np.array([0.5]))
    on = OracleNodes(gn)
    on.ca, on.ce = ca, np.zeros(1)
    kw = dict(strategy="grid", latent_marginal_indices=[0, 5], predictor_marginal_indices=[3],
              lik=(fam, link, y, oa, ob))
    rg = I.integrate(gn, np.array([0.5]), **kw)
    ro = I.integrate(on, np.array([0.5]), **kw)
    _compare(rg, ro, record_error, "areal Poisson, grid")
    # the public entry point gives the same result, and the constraint holds at every node
    res = I.run_inla_inference([0.5], build_prior, log_prior, obs, strategy="grid", a=A,
                               constraints_a=ca.ravel().tolist(), constraints_e=[0.0],
                               latent_marginal_indices=[0, 5], predictor_marginal_indices=[3], capacity=8)
    assert np.allclose(res.mode, rg["mode"], rtol=0, atol=1e-12)
    assert abs(res.marginal_log_lik - rg["marginal_log_lik"]) < 1e-9
    assert np.max(np.abs(ca @ rg["x_nodes"].T)) < 1e-10
    assert res.posterior_precision.shape == (A.shape[1], A.shape[1])
    assert len(res.cpo) == len(obs) and len(res.marginals_latent) == 2


def test_front_end_gaussian_free_precision_ccd_matches_oracle_nodes(record_error):
    from inla_gpu import inference as I
    from inla_gpu import workloads as W

    wl = W.spde_gaussian_workload(nx=10, n_obs=300, seed=21)
    g = wl.gmrf

    def build_prior(theta):
        c, _ = W.spde_coefs(theta[1], theta[2])
        return g.csc(c)

    def log_prior(theta):
        return wl.log_prior(theta)[0]

    obs = [{"family": "gaussian", "y": v, "precision": 1.0, "prec_theta": 0} for v in wl.y]
    fam, link, y, oa, ob, slot = I.parse_obs(obs, gaussian_free_prec=True)
    assert slot == 0
    th0 = wl.theta_true + np.array([0.3, -0.2, 0.2])
    gn = I.GpuNodes(build_prior, log_prior, 3, fam, link, y, oa, ob, slot, A=wl.A, capacity=16, theta0=th0)
    on = OracleNodes(gn)
    kw = dict(strategy="ccd", latent_marginal_indices=[7], lik=(fam, link, y, oa, ob))
    rg = I.integrate(gn, th0, **kw)
    ro = I.integrate(on, th0, **kw)
    _compare(rg, ro, record_error, "SPDE Gaussian, CCD")
    # the adjoint-gradient search reaches the same mode
    rb = I.integrate(gn, th0, mode_search="bfgs", **kw)
    assert np.max(np.abs(rb["mode"] - rg["mode"])) < 2e-3


@pytest.mark.parametrize("family", ["poisson", "gaussian"])
def test_values_mode_matches_coefficient_mode(family):
    """gmrf_eval_args.qvals / dqvals: Q(theta_b) and dQ/dtheta given as values on the
    pattern give the coefficient form's mlik, mode, variances and gradient (the
    contractions run over the derivative matrices instead of the components)."""
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    if family == "poisson":
        wl = W.spde_poisson_workload(nx=12, n_obs=800, seed=4)
        th = wl.theta_true[None] + np.array([[0.0, 0.0], [0.2, -0.1], [-0.1, 0.2]])
        c, dc, lp, dl = wl.params(th)
    else:
        wl = W.spde_gaussian_workload(nx=12, n_obs=500, seed=4)
        th = wl.theta_true[None] + np.array([[0.0, 0.0, 0.0], [0.2, -0.1, 0.1]])
        c, dc, lp, dl = wl.params(th)
    g = wl.gmrf
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family=family, A=wl.A, capacity=3)
    rc = model.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    qv = c @ g.components
    dq = np.einsum("bmk,ke->bme", dc, g.components)
    rv = model.eval(qvals=qv, dqvals=dq, lpar=lp, dlog_lpar=dl)
    assert (rv["status"] == 0).all()
    for k in ("mlik", "x", "var", "grad"):
        a, b = rv[k].cpu().numpy(), rc[k].cpu().numpy()
        assert np.max(np.abs(a - b)) <= 1e-11 * max(np.max(np.abs(b)), 1.0), (k, np.max(np.abs(a - b)))
