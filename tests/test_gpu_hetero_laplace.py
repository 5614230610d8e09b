"""Heterogeneous Laplace batches: independent BYM2 + Poisson models of uneven size as
one block-diagonal model (LaplaceModel(blocks=...)), Newton, mlik, variances and the
adjoint gradient per (theta-point, block), each block against the oracle run on that
model alone (crates/inla_stats/src/inference.rs:776-975 per model).

Bars as for single models: mlik 1e-8 absolute, mode 1e-10 relative, marginal
variances 1e-9 relative, the gradient against central differences (step 1e-4) of the
oracle's mlik at 1e-5 relative.
"""

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MLIK_ATOL = 1e-8
MODE_RTOL = 1e-10
VAR_RTOL = 1e-9


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def small_models(sizes, seed=5, tau=1.0, phi=0.5, intercept=1.0):
    """BYM2 + Poisson models on Delaunay-neighbour graphs of the given sizes (the
    scale_model factor from the dense generalised inverse)."""
    import scipy.sparse as sp

    from inla_gpu import workloads as W

    rng = np.random.default_rng(seed)
    out = []
    for s in sizes:
        adj = W.delaunay_adjacency(int(s), rng)
        R = W.besag_structure(adj)
        m = W.Bym2Model(adj, R, W.scale_factor_dense(R), np.zeros(R.shape[0]))
        Q = sp.csc_matrix((m.values(tau, phi), m.indices, m.indptr), shape=(m.n, m.n))
        u = W.sample_gmrf(Q, rng)
        m.y = rng.poisson(np.exp(intercept + u)).astype(float)
        out.append(m)
    return out


@pytest.fixture(scope="module")
def batch():
    from inla_gpu import workloads as W

    return W.bym2_poisson_batch(models=small_models([60, 310, 45, 180, 400, 120]))


def _oracle(wl, q, th, keep_post=False):
    return orc.find_mode(wl.block_prior(q, th), "poisson", "log", wl.block_y(q), 1.0, A=wl.block_A(q),
                         keep_post=keep_post)


def test_blocks_match_oracle_per_model(batch, record_error):
    from inla_gpu.laplace import LaplaceModel

    wl = batch
    k = wl.k
    rng = np.random.default_rng(3)
    thetas = rng.normal(0.0, 0.6, size=(3, k, 2))
    model = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=4)
    c, dc, _, _ = wl.params(thetas)
    r = model.eval(c, dcoefs=dc)
    assert r["mlik"].shape == (3, k) and r["grad"].shape == (3, k, 2) and r["status"].shape == (3, k)
    assert (r["status"] == 0).all(), r["status"]
    ml, x, var, gr = (r[key].cpu().numpy() for key in ("mlik", "x", "var", "grad"))
    bp = wl.block_ptr
    h = 1e-4
    for b in range(3):
        for q in range(k):
            o = _oracle(wl, q, thetas[b, q], keep_post=True)
            e_ml = record_error("hetero mlik", abs(ml[b, q] - o.mlik), MLIK_ATOL, "abs")
            e_x = record_error("hetero mode", _rel(x[b, bp[q]:bp[q + 1]], o.x), MODE_RTOL)
            e_v = record_error("hetero variances", _rel(var[b, bp[q]:bp[q + 1]], o.post_factor().takahashi_diag()),
                               VAR_RTOL)
            assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL, (b, q, e_ml, e_x, e_v)
    # the gradient of every block at the first point against central differences
    for q in range(k):
        for l in range(2):
            tp, tm = thetas[0, q].copy(), thetas[0, q].copy()
            tp[l] += h
            tm[l] -= h
            fd = (_oracle(wl, q, tp).mlik - _oracle(wl, q, tm).mlik) / (2 * h)
            e_g = record_error("hetero gradient vs FD", abs(gr[0, q, l] - fd) / max(abs(fd), 1.0), 1e-5)
            assert e_g < 1e-5, (q, l, gr[0, q, l], fd)


def test_blocks_equal_separate_models(batch):
    """Each block of the union evaluates as its model alone does (the union changes
    nothing but the launch sequence): mlik, mode, gradient and Newton iterations."""
    from inla_gpu.laplace import LaplaceModel

    wl = batch
    k = wl.k
    thetas = np.random.default_rng(4).normal(0.0, 0.6, size=(2, k, 2))
    union = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=2)
    c, dc, _, _ = wl.params(thetas)
    ru = union.eval(c, dcoefs=dc)
    bp = wl.block_ptr
    for q in range(k):
        e0, e1 = wl.indptr[bp[q]], wl.indptr[bp[q + 1]]
        single = LaplaceModel(bp[q + 1] - bp[q], wl.indptr[bp[q]:bp[q + 1] + 1] - e0, wl.indices[e0:e1] - bp[q],
                              wl.components[:, e0:e1], wl.block_y(q), family="poisson", A=wl.block_A(q),
                              capacity=2)
        rs = single.eval(c[:, q], dcoefs=dc[:, q])
        assert np.array_equal(rs["n_iter"], ru["n_iter"][:, q])
        assert np.allclose(rs["mlik"].cpu().numpy(), ru["mlik"][:, q].cpu().numpy(), rtol=0, atol=1e-10)
        assert np.allclose(rs["x"].cpu().numpy(), ru["x"][:, bp[q]:bp[q + 1]].cpu().numpy(), rtol=1e-12, atol=1e-13)
        assert np.allclose(rs["grad"].cpu().numpy(), ru["grad"][:, q].cpu().numpy(), rtol=1e-9, atol=1e-10)


def test_failing_block_is_isolated(batch):
    """A block whose coefficients are not finite fails alone; the other blocks of the
    same point keep their results."""
    from inla_gpu.laplace import LaplaceModel

    wl = batch
    k = wl.k
    thetas = np.zeros((2, k, 2))
    model = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=2)
    c, dc, _, _ = wl.params(thetas)
    good = model.eval(c, dcoefs=dc)
    c_bad = c.copy()
    c_bad[1, 2, 0] = np.nan
    r = model.eval(c_bad, dcoefs=dc)
    assert r["status"][1, 2] != 0
    ok = np.ones((2, k), dtype=bool)
    ok[1, 2] = False
    assert (r["status"][ok] == 0).all()
    ml, mg = r["mlik"].cpu().numpy(), good["mlik"].cpu().numpy()
    assert np.allclose(ml[ok], mg[ok], rtol=0, atol=1e-10)
    assert ml[1, 2] == -np.inf


def test_block_arguments_are_checked(batch):
    from inla_gpu.laplace import LaplaceModel
    from inla_gpu.solver import InvalidArgument

    wl = batch
    bp = wl.block_ptr.copy()
    bad = bp.copy()
    bad[2] = bad[1]
    with pytest.raises(InvalidArgument):
        LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A, blocks=bad)
    # an observation loading two blocks
    A2 = wl.A.tolil()
    A2[0, bp[1]] = 1.0
    with pytest.raises(Exception):
        LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=A2.tocsc(),
                     blocks=bp)
    with pytest.raises(InvalidArgument):
        LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A, blocks=bp,
                     constraints=(np.ones((1, wl.n)), np.zeros(1)))


def test_full_size_64_models_match_oracle_per_model(record_error):
    """The full heterogeneous configuration (64 BYM2 + Poisson models with an intercept
    each, Delaunay-neighbour graphs of sizes uniform in [500, 20000], n = 592,548) as
    one block-diagonal model at three theta-points per model; every (point, model)
    pair's mlik, mode and marginal variances against the oracle run on that model
    alone, the oracle's calls spread over the host cores."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    import torch

    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.bym2_poisson_batch(k=64)
    k = len(wl.models)
    assert k == 64
    model = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=3)
    offs = np.array([[0.0, 0.0], [0.4, -0.3], [-0.3, 0.5]])
    thetas = np.broadcast_to(offs[:, None, :], (3, k, 2)).copy()
    c, _, _, _ = wl.params(thetas)
    r = model.eval(torch.as_tensor(c, device="cuda"))
    assert (r["status"] == 0).all()
    ml, x, var = r["mlik"].cpu().numpy(), r["x"].cpu().numpy(), r["var"].cpu().numpy()
    bp = wl.block_ptr

    def ref(job):
        b, q = job
        o = orc.find_mode(wl.block_prior(q, thetas[b, q]), "poisson", "log", wl.block_y(q), 1.0,
                          A=wl.block_A(q), keep_post=True)
        return b, q, o.mlik, o.x, o.post_factor().takahashi_diag()

    jobs = [(b, q) for b in range(3) for q in range(k)]
    with ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0))))) as ex:
        refs = list(ex.map(ref, jobs))
    e_ml = e_x = e_v = 0.0
    for b, q, o_ml, o_x, o_v in refs:
        e_ml = max(e_ml, abs(ml[b, q] - o_ml))
        e_x = max(e_x, _rel(x[b, bp[q]:bp[q + 1]], o_x))
        e_v = max(e_v, _rel(var[b, bp[q]:bp[q + 1]], o_v))
    record_error("64 models x 3 points: mlik", e_ml, MLIK_ATOL, "abs")
    record_error("64 models x 3 points: mode", e_x, MODE_RTOL)
    record_error("64 models x 3 points: variances", e_v, VAR_RTOL)
    assert e_ml < MLIK_ATOL and e_x < MODE_RTOL and e_v < VAR_RTOL
