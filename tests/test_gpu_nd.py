"""The nested-dissection and automatic orderings on the device against the oracle:
batched LDL^T, solve, selected inversion (diagonal and pattern entries) of matrices
ordered by nested dissection, and Laplace evaluations of models analysed with
order="nd" / "auto" (Gaussian at full headline size, Poisson with an intercept, binomial
and negative binomial at test size): the oracle orders by its own minimum degree, so
these checks cover the factor under a different elimination order. Tolerances: log
determinant and solve 1e-10 relative, selected inverse 1e-9 relative, mlik 1e-8 absolute,
mode 1e-10 relative, variances 1e-9 relative."""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LOGDET_RTOL = 1e-10
SOLVE_RTOL = 1e-10
SELINV_RTOL = 1e-9
MLIK_ATOL = 1e-8
MODE_RTOL = 1e-10
VAR_RTOL = 1e-9


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pool():
    return ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0)))))


def _spde_post(nx, seed=0):
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(nx, nx)
    rng = np.random.default_rng(seed)
    m = int(2.5 * g.n)
    A = W.lattice_projector(nx, nx, rng.uniform(size=(m, 2)))
    p = sp.csc_matrix(g.csc([1.0, 1.0, 1.0]) + A.T @ A)
    p.sort_indices()
    return p


def _besag_ar1(nr, T):
    from inla_gpu import workloads as W

    R = W.besag_structure(W.rook_adjacency(nr, nr))
    Ra = sp.diags([np.full(T - 1, -0.7), np.r_[1.0, np.full(T - 2, 1.49), 1.0], np.full(T - 1, -0.7)], [-1, 0, 1])
    q = sp.csc_matrix(sp.kron(Ra, R + 0.05 * sp.identity(nr * nr)) + 0.1 * sp.identity(nr * nr * T))
    q.sort_indices()
    return q


CASES = {
    "spde_60": lambda: _spde_post(60),
    "besag_ar1_16x16x8": lambda: _besag_ar1(16, 8),
    "random_400": lambda: sp.csc_matrix(M.random_spd(400, 0.01, 5)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_nd_batch_matches_oracle_per_matrix(name, record_error):
    import torch

    from inla_gpu import _lib
    from inla_gpu.solver import Batch, Symbolic

    q0 = CASES[name]()
    q0.sort_indices()
    n = q0.shape[0]
    sym = Symbolic(n, q0.indptr, q0.indices, order=_lib.ORDER_ND)
    assert sym.stats()["order_kind"] == 5
    rng = np.random.default_rng(3)
    cols = np.repeat(np.arange(n), np.diff(q0.indptr))
    diag = q0.indices == cols
    B = 3
    vals = np.stack([q0.data * (1.0 + 0.1 * b) + np.where(diag, 0.5 * b, 0.0) for b in range(B)])
    bt = Batch(sym, B)
    bt.factorize(torch.from_numpy(vals).cuda())
    ld, st = bt.logdet()
    assert (st == 0).all()
    rhs = rng.standard_normal((B, n))
    x = bt.solve(torch.from_numpy(rhs).cuda()).cpu().numpy()
    bt.selinv()
    d = bt.selinv_diag().cpu().numpy()
    sig = bt.selinv_pattern().cpu().numpy()
    for b in range(B):
        q = sp.csc_matrix((vals[b], q0.indices, q0.indptr), shape=(n, n))
        ref = orc.Factor(q)
        e_ld = record_error(f"logdet ({name})", abs(ld[b] - ref.log_abs_det()) / abs(ref.log_abs_det()), LOGDET_RTOL)
        e_x = record_error(f"solve ({name})", _rel(x[b], ref.solve(rhs[b])), SOLVE_RTOL)
        e_d = record_error(f"selinv diagonal ({name})", _rel(d[b], ref.takahashi_diag()), SELINV_RTOL)
        e_s = record_error(f"selinv pattern ({name})", _rel(sig[b], ref.selinv_entries(q0.indices, cols)), SELINV_RTOL)
        assert e_ld < LOGDET_RTOL and e_x < SOLVE_RTOL and e_d < SELINV_RTOL and e_s < SELINV_RTOL


def test_auto_model_full_size_gaussian_matches_oracle(record_error):
    """The headline model (142 x 142 SPDE lattice, 50,000 Gaussian observations, 15
    theta-points) analysed with order="auto" (which picks nested dissection here):
    every point's mlik, mode, variances and gradient against the chooser-ordered model,
    and three points against the oracle directly."""
    from inla_gpu import ccd
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload()
    g = wl.gmrf
    z, _ = ccd.ccd_design(3, 1.0)
    thetas = wl.theta_true[None] + 0.1 * z
    coefs, dcoefs, lpar, dlog = wl.params(thetas)
    outs = {}
    for order in ("chooser", "auto"):
        model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A,
                             capacity=15, order=order)
        r = model.eval(coefs, lpar=lpar, dcoefs=dcoefs, dlog_lpar=dlog)
        assert (r["status"] == 0).all(), (order, r["status"])
        outs[order] = {k: r[k].cpu().numpy() for k in ("mlik", "x", "var", "grad")}
    a, c = outs["auto"], outs["chooser"]
    e = record_error("auto vs chooser mlik (15 points)", float(np.max(np.abs(a["mlik"] - c["mlik"]))), MLIK_ATOL, "abs")
    assert e < MLIK_ATOL
    for k, tol in (("x", MODE_RTOL), ("var", VAR_RTOL), ("grad", 1e-9)):
        e = record_error(f"auto vs chooser {k} (15 points)", _rel(a[k], c[k]), tol)
        assert e < tol, k

    def ref(b):
        o = orc.find_mode(g.csc(coefs[b]), "gaussian", "identity", wl.y, np.full(len(wl.y), lpar[b]), A=wl.A,
                          keep_post=True)
        return o.mlik, o.x, o.post_factor().takahashi_diag()

    with _pool() as ex:
        refs = list(ex.map(ref, (0, 5, 14)))
    for b, (ml, x, v) in zip((0, 5, 14), refs):
        e_ml = record_error("auto vs oracle mlik", abs(a["mlik"][b] - ml), MLIK_ATOL, "abs")
        e_x = record_error("auto vs oracle mode", _rel(a["x"][b], x), MODE_RTOL)
        e_v = record_error("auto vs oracle variances", _rel(a["var"][b], v), VAR_RTOL)
        assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (b, e_ml, e_x, e_v)


@pytest.mark.parametrize("family", ["poisson", "binomial", "nbinomial"])
def test_nd_model_families_match_oracle(family, record_error):
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload(nx=24, n_obs=1500, seed=11)
    g = wl.gmrf
    rng = np.random.default_rng(4)
    y = np.asarray(wl.y, dtype=np.float64)
    pa = pb = None
    link = "log"
    if family == "binomial":
        pa = np.full(len(y), 12.0)
        y = np.minimum(y, pa)
        link = "logit"
    elif family == "nbinomial":
        pb = np.full(len(y), 4.0)
    thetas = wl.theta_true[None] + np.array([[0.0, 0.0], [0.2, -0.1], [-0.15, 0.2]])
    coefs, dcoefs, _, _ = wl.params(thetas)
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, y, family=family, A=wl.A, obs_a=pa, obs_b=pb,
                         capacity=3, order="nd")
    r = model.eval(coefs, dcoefs=dcoefs)
    assert (r["status"] == 0).all(), r["status"]
    for b in range(3):
        o = orc.find_mode(g.csc(coefs[b]), family, link, y, pa if pa is not None else 1.0, pb=pb, A=wl.A,
                          keep_post=True)
        e_ml = record_error(f"nd {family} mlik", abs(r["mlik"][b].item() - o.mlik), MLIK_ATOL, "abs")
        e_x = record_error(f"nd {family} mode", _rel(r["x"][b].cpu().numpy(), o.x), MODE_RTOL)
        e_v = record_error(f"nd {family} variances", _rel(r["var"][b].cpu().numpy(), o.post_factor().takahashi_diag()),
                           VAR_RTOL)
        assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (b, e_ml, e_x, e_v)
