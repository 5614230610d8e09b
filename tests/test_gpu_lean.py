"""Lean evaluation (gmrf_model_desc.lean): the prior factorised (and inverted, for the
gradient) alone before the Newton loop in the posterior's batch slots, so the device
holds one factor per point instead of two. Results must equal the standard evaluation
(mlik 1e-10 absolute, mode / variances / gradient 1e-12 relative) for every path the
standard evaluation has: general likelihoods, the Gaussian evaluation (captured graph in
the standard form), the low-rank constraint correction and heterogeneous blocks; and
the oracle at the usual bars. (The Gaussian comparison allows the rounding of the
standard form's on-device finish.)
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KAPPA = 3.059023205018258e6  # constraints.rs:268


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _same(r_lean, r_std, keys=("x", "var", "grad"), mlik_tol=1e-10, rtol=1e-12):
    assert np.array_equal(r_lean["status"], r_std["status"])
    assert np.array_equal(r_lean["n_iter"], r_std["n_iter"])
    ml, ms = r_lean["mlik"].cpu().numpy(), r_std["mlik"].cpu().numpy()
    assert np.max(np.abs(ml - ms)) <= mlik_tol, np.max(np.abs(ml - ms))
    for k in keys:
        if k in r_std:
            e = _rel(r_lean[k].cpu().numpy(), r_std[k].cpu().numpy())
            assert e <= rtol, (k, e)


def test_lean_negbin_space_time_matches_standard_and_oracle(record_error):
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_rw2_workload(nx=8, T=4, n_obs=400, seed=17)
    g = wl.gmrf
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0], [0.2, -0.15], [-0.1, 0.1]])
    kw = dict(family="nbinomial", A=wl.A, obs_a=wl.obs_a, obs_b=wl.obs_b, capacity=3)
    std = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, **kw)
    lean = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, lean=True, **kw)
    c, dc, _, _ = wl.params(thetas)
    rs, rl = std.eval(c, dcoefs=dc), lean.eval(c, dcoefs=dc)
    assert (rl["status"] == 0).all()
    _same(rl, rs)
    for b in range(len(thetas)):
        ref = orc.find_mode(g.csc(c[b]), "nbinomial", "log", wl.y, wl.obs_a, pb=wl.obs_b, A=wl.A,
                            keep_post=True)
        e_ml = record_error("mlik (lean)", abs(rl["mlik"][b].item() - ref.mlik), 1e-8, "abs")
        e_x = record_error("mode (lean)", _rel(rl["x"][b].cpu().numpy(), ref.x), 1e-10)
        e_v = record_error("variances (lean)", _rel(rl["var"][b].cpu().numpy(), ref.post_factor().takahashi_diag()),
                           1e-9)
        assert e_ml < 1e-8 and e_x < 1e-10 and e_v < 1e-9
    # the variances alone (no gradient: no prior inversion) and the mlik alone
    _same(lean.eval(c), std.eval(c), keys=("x", "var"))
    _same(lean.eval(c, want_var=False), std.eval(c, want_var=False), keys=("x",))


def test_lean_gaussian_matches_graph_evaluation():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload(nx=16, n_obs=900, seed=3)
    g = wl.gmrf
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0, 0.0], [0.2, -0.1, 0.1]])
    c, dc, lp, dl = wl.params(thetas)
    std = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A, capacity=2)
    lean = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A, capacity=2,
                        lean=True)
    rs = std.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    rl = lean.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    assert (rl["status"] == 0).all()
    # the standard form finishes mlik and gradient on the device, the general path on the
    # host: the same terms summed in another order
    _same(rl, rs, mlik_tol=1e-12 * float(np.max(np.abs(rs["mlik"].cpu().numpy()))), rtol=1e-11)


def test_lean_lowrank_constraints_match_standard():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    nr, T = 6, 4
    wl = W.besag_ar1_workload(nr=nr, T=T, seed=13)
    ca_raw, ce = wl.constraints
    ca = ca_raw / math.sqrt(nr * nr)
    g = wl.gmrf
    kw = dict(family="binomial", obs_a=wl.obs_a, constraints=(ca, ce), capacity=2, constraint_kappa=KAPPA)
    std = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, **kw)
    lean = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, lean=True, **kw)
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0], [0.3, -0.2]])
    c, dc, _, _ = wl.params(thetas)
    rs, rl = std.eval(c, dcoefs=dc), lean.eval(c, dcoefs=dc)
    assert (rl["status"] == 0).all()
    _same(rl, rs)
    assert np.max(np.abs(ca @ rl["x"].cpu().numpy().T)) < 1e-12


def test_lean_blocks_match_standard():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    rng = np.random.default_rng(8)
    models = []
    for s in (50, 130, 75):
        adj = W.delaunay_adjacency(s, rng)
        R = W.besag_structure(adj)
        m = W.Bym2Model(adj, R, W.scale_factor_dense(R), np.zeros(s))
        Q = sp.csc_matrix((m.values(1.0, 0.5), m.indices, m.indptr), shape=(s, s))
        m.y = rng.poisson(np.exp(1.0 + W.sample_gmrf(Q, rng))).astype(float)
        models.append(m)
    wl = W.bym2_poisson_batch(models=models)
    thetas = rng.normal(0.0, 0.5, size=(2, wl.k, 2))
    c, dc, _, _ = wl.params(thetas)
    kw = dict(family="poisson", A=wl.A, blocks=wl.block_ptr, capacity=2)
    std = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, **kw)
    lean = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, lean=True, **kw)
    rs, rl = std.eval(c, dcoefs=dc), lean.eval(c, dcoefs=dc)
    assert (rl["status"] == 0).all()
    _same(rl, rs)
