"""HIP path vs the oracle: factorisation, log|Q|, solves, selected inversion.

Tolerances follow the parity bar: log|Q| 1e-10 relative, solves 1e-10 relative
(in the max norm), selected-inverse entries 1e-9 relative.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LOGDET_RTOL = 1e-10
SOLVE_RTOL = 1e-10
SELINV_RTOL = 1e-9


def _solver():
    from inla_gpu.solver import HipSolver

    return HipSolver()


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def spde_q(nx, obs_per_vertex=2.5, seed=0, theta=(0.0, 1.5)):
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(nx, nx)
    c, _ = W.spde_coefs(*theta)
    rng = np.random.default_rng(seed)
    m = int(obs_per_vertex * g.n)
    A = W.lattice_projector(nx, nx, rng.uniform(size=(m, 2)))
    q = g.csc(c) + sp.csc_matrix(A.T @ sp.diags(np.full(m, 100.0)) @ A)
    q = sp.csc_matrix(q)
    q.sum_duplicates()
    q.sort_indices()
    return q


def check_against_oracle(q, seed=1):
    q = sp.csc_matrix(q)
    n = q.shape[0]
    rng = np.random.default_rng(seed)
    s = _solver()
    s.factorize(q)
    ref = orc.Factor(q)
    assert ref.status == 0
    ld, ld_ref = s.log_abs_det(), ref.log_abs_det()
    assert abs(ld - ld_ref) <= LOGDET_RTOL * max(abs(ld_ref), 1.0), (ld, ld_ref)
    b = rng.standard_normal(n)
    x = s.solve(b)
    assert _rel(x, ref.solve(b)) < SOLVE_RTOL
    d = s.diag_inv()
    assert _rel(d, ref.takahashi_diag()) < SELINV_RTOL
    return s, ref


def test_mfma_tile_layout():
    from inla_gpu.solver import probe_mma

    rng = np.random.default_rng(3)
    for K in (4, 17, 32, 64):
        A = rng.standard_normal((64, K))
        B = rng.standard_normal((64, K))
        T = probe_mma(A, B)
        assert np.max(np.abs(T - A @ B.T)) < 1e-12 * K


def test_identity_solver_known_answer():
    # crates/inla_math/src/solver.rs:106-121
    s = _solver()
    s.factorize(sp.identity(3, format="csc") * 2.0)
    assert np.array_equal(s.solve(np.array([2.0, 4.0, 6.0])), [1.0, 2.0, 3.0])
    assert np.array_equal(s.diag_inv(), [0.5, 0.5, 0.5])
    assert abs(s.log_abs_det() - 3 * np.log(2.0)) < 1e-10


@pytest.mark.parametrize("q", [M.ar1_q(12, 0.7, 2.5), M.ar1_q(3, 0.2, 1.0),
                               M.besag_ridge_q(M.FOUR_CYCLE_CHORD, 1.5, 0.25),
                               M.besag_ridge_q(M.PATH6, 1.0, 0.1)],
                         ids=["ar1_12", "ar1_3", "besag4", "path6"])
def test_reference_takahashi_cases(q):
    # sparse_ldlt.rs:613-665: Takahashi vs dense inverse / blocked solves
    s, ref = check_against_oracle(q)
    dense = np.diag(np.linalg.inv(q.toarray()))
    assert np.max(np.abs(s.diag_inv() - dense)) < 1e-9


def test_ar1_closed_form_logdet():
    n, rho, tau = 500, 0.9, 3.0
    s = _solver()
    s.factorize(M.ar1_q(n, rho, tau))
    exact = n * np.log(tau) + np.log(1 - rho * rho)
    assert abs(s.log_abs_det() - exact) < 1e-10 * abs(exact)


@pytest.mark.parametrize("n,density,seed", [(60, 0.1, 0), (300, 0.02, 1), (400, 0.3, 2)])
def test_random_spd(n, density, seed):
    check_against_oracle(M.random_spd(n, density, seed), seed)


def test_dense_front_902():
    # the reference factors C1 as a dense 902 x 902 matrix after kappa A^T A augmentation
    rng = np.random.default_rng(5)
    a = rng.standard_normal((902, 902)) / 30
    q = sp.csc_matrix(a @ a.T + np.eye(902) * 2.0)
    check_against_oracle(q)


def test_spde_small():
    check_against_oracle(spde_q(30))


def test_spde_full_size_logdet_solve():
    # n = 20,164 SPDE lattice + A^T W A; diag via the oracle Takahashi is slow, so
    # compare the diagonal on a strided sample of selected-inverse entries.
    q = spde_q(142, obs_per_vertex=2.48)
    n = q.shape[0]
    s = _solver()
    s.factorize(q)
    ref = orc.Factor(q)
    ld_ref = ref.log_abs_det()
    assert abs(s.log_abs_det() - ld_ref) <= LOGDET_RTOL * abs(ld_ref)
    b = np.random.default_rng(0).standard_normal(n)
    assert _rel(s.solve(b), ref.solve(b)) < SOLVE_RTOL
    d = s.diag_inv()
    idx = np.arange(0, n, 97)
    cols = np.zeros((n, len(idx)))
    cols[idx, np.arange(len(idx))] = 1.0
    # diag entries through oracle solves (Q^-1 e_i)_i
    dref = np.array([ref.solve(cols[:, k])[i] for k, i in enumerate(idx)])
    assert _rel(d[idx], dref) < SELINV_RTOL


def test_singular_and_not_pd_errors():
    from inla_gpu.solver import NotPositiveDefinite, Singular, SolveFailure

    s = _solver()
    with pytest.raises(SolveFailure):
        s.solve(np.ones(3))
    with pytest.raises(SolveFailure):
        s.diag_inv()
    # exact zero pivot: the numeric phase fails (sparse_ldlt.rs:230 -> NotPositiveDefinite)
    with pytest.raises(NotPositiveDefinite):
        s.factorize(sp.csc_matrix(np.diag([1.0, 0.0, 2.0])))
    # tiny pivot: |D| < 1e-14 -> Singular (sparse_ldlt.rs:118-119)
    with pytest.raises(Singular):
        s.factorize(sp.csc_matrix(np.diag([1.0, 1e-16, 2.0])))
    # indefinite but regular: LDL^T succeeds, log|Q| = sum log|D|
    s.factorize(sp.csc_matrix(np.diag([1.0, -4.0, 2.0])))
    assert abs(s.log_abs_det() - np.log(8.0)) < 1e-14


def test_failed_factorize_keeps_previous_factor():
    from inla_gpu.solver import Singular

    s = _solver()
    q = M.ar1_q(20, 0.5, 2.0)
    s.factorize(q)
    b = np.arange(20.0)
    x0 = s.solve(b)
    with pytest.raises(Singular):
        s.factorize(sp.csc_matrix(np.diag(np.r_[np.ones(19), 1e-20])))
    assert np.array_equal(s.solve(b), x0)


def test_pattern_reuse_and_refactorize():
    q1 = M.ar1_q(50, 0.3, 1.0)
    q2 = M.ar1_q(50, 0.8, 4.0)
    s = _solver()
    s.factorize(q1)
    s.factorize(q2)
    ref = orc.Factor(q2)
    assert abs(s.log_abs_det() - ref.log_abs_det()) < 1e-10 * abs(ref.log_abs_det())


def test_batch_matches_oracle_per_matrix():
    import torch

    from inla_gpu import workloads as W
    from inla_gpu.solver import Batch, Symbolic

    g, _, _ = W.spde_components(40, 40)
    thetas = [(-1.0, 1.0), (0.0, 1.5), (0.5, 2.0), (1.0, 2.2)]
    vals = np.stack([g.values(W.spde_coefs(*t)[0]) + 1e-3 * np.where(
        g.indices == np.repeat(np.arange(g.n), np.diff(g.indptr)), 1.0, 0.0) for t in thetas])
    sym = Symbolic(g.n, g.indptr, g.indices)
    bt = Batch(sym, len(thetas))
    dv = torch.from_numpy(vals).cuda()
    bt.factorize(dv)
    ld, st = bt.logdet()
    assert (st == 0).all()
    rng = np.random.default_rng(2)
    rhs = rng.standard_normal((len(thetas), g.n))
    x = bt.solve(torch.from_numpy(rhs).cuda()).cpu().numpy()
    bt.selinv()
    diag = bt.selinv_diag().cpu().numpy()
    sig = bt.selinv_pattern().cpu().numpy()
    cols = np.repeat(np.arange(g.n), np.diff(g.indptr))
    for b in range(len(thetas)):
        q = sp.csc_matrix((vals[b], g.indices, g.indptr), shape=(g.n, g.n))
        ref = orc.Factor(q)
        assert abs(ld[b] - ref.log_abs_det()) <= LOGDET_RTOL * abs(ref.log_abs_det())
        assert _rel(x[b], ref.solve(rhs[b])) < SOLVE_RTOL
        assert _rel(diag[b], ref.takahashi_diag()) < SELINV_RTOL
        assert _rel(sig[b], ref.selinv_entries(g.indices, cols)) < SELINV_RTOL


@pytest.mark.parametrize("case", ["spde", "dense"])
def test_solve_multi_equals_single_solves(case):
    """Several right-hand sides per matrix in one launch sequence run the same
    arithmetic per vector as one-vector solves: the results are bit-identical."""
    import torch

    from inla_gpu import workloads as W
    from inla_gpu.solver import Batch, Symbolic

    if case == "spde":
        g, _, _ = W.spde_components(40, 40)
        q = sp.csc_matrix(g.csc(W.spde_coefs(0.0, 1.5)[0]) + 1e-3 * sp.identity(g.n))
    else:
        q = sp.csc_matrix(M.random_spd(400, 0.3, 2))  # large fronts
    q.sort_indices()
    n = q.shape[0]
    bt = Batch(Symbolic.from_csc(q), 2)
    bt.factorize(torch.from_numpy(np.stack([q.data, 3.0 * q.data])).cuda())
    rng = np.random.default_rng(4)
    rhs = torch.from_numpy(rng.standard_normal((2, 5, n))).cuda()
    xm = bt.solve_multi(rhs).cpu().numpy()
    for j in range(5):
        xs = bt.solve(rhs[:, j, :].contiguous()).cpu().numpy()
        assert np.array_equal(xm[:, j, :], xs)
    ref = orc.Factor(q)
    assert _rel(xm[0, 3], ref.solve(rhs[0, 3].cpu().numpy())) < SOLVE_RTOL


@pytest.mark.parametrize("n,bw", [(3000, 1100), (4500, 2200)])
def test_tall_fronts_split_solve_and_selinv(n, bw):
    """A banded matrix under the natural order: a chain of fronts with more than
    1024 rows below the supernode, so the solves take the split path (triangle per
    front, row-parallel GEMV launches for the rows below); at bandwidth 2200 the
    rows below a panel exceed 2048 and Sigma[P,P] is formed from partial products.
    Dense numpy references."""
    import torch

    from inla_gpu import _lib as LL
    from inla_gpu.solver import Batch, Symbolic

    rng = np.random.default_rng(0)
    rows, cols = [], []
    for d in (1, 2, bw):
        rows += list(range(d, n))
        cols += list(range(0, n - d))
    a = sp.csc_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    q = sp.csc_matrix(a + a.T + sp.diags(np.abs(a).sum(0).A1 + np.abs(a).sum(1).A1 + 1.0))
    q.sort_indices()
    sym = Symbolic.from_csc(q, order=LL.ORDER_NATURAL)
    assert sym.stats()["max_front"] > bw
    bt = Batch(sym, 2)
    bt.factorize(torch.from_numpy(np.stack([q.data, 3.0 * q.data])).cuda())
    ld, st = bt.logdet()
    assert (st == 0).all()
    qd = q.toarray()
    ref_ld = np.linalg.slogdet(qd)[1]
    assert abs(ld[0] - ref_ld) < LOGDET_RTOL * abs(ref_ld)
    assert abs(ld[1] - (ref_ld + n * np.log(3.0))) < LOGDET_RTOL * abs(ref_ld)
    rhs = rng.standard_normal((2, 3, n))
    xm = bt.solve_multi(torch.from_numpy(rhs).cuda()).cpu().numpy()
    for j in range(3):
        xs = np.linalg.solve(qd, rhs[0, j])
        assert _rel(xm[0, j], xs) < SOLVE_RTOL
        assert _rel(xm[1, j], np.linalg.solve(qd, rhs[1, j]) / 3.0) < SOLVE_RTOL
    bt.selinv()
    dg = bt.selinv_diag().cpu().numpy()
    di = np.diag(np.linalg.inv(qd))
    assert _rel(dg[0], di) < SELINV_RTOL
    assert _rel(dg[1], di / 3.0) < SELINV_RTOL


@pytest.mark.parametrize("n", [1, 7, 64, 150])
def test_factorize_dense_matches_numpy(n):
    # LdltBackend::factorize_dense (backend.rs:59-69, dense_faer.rs:25-75): the lower
    # triangle of a row-major n x n matrix; one dense front (n = 7 and 64: small-front
    # kernels, 150: the tile-parallel path). log|A|, a solve and diag(A^-1) against
    # numpy on the same matrix; the upper triangle holds garbage that must not be read.
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    full = G @ G.T / n + np.eye(n)
    a_in = full.copy()
    a_in[np.triu_indices(n, 1)] = 1e300
    s = _solver()
    s.factorize_dense(a_in)
    _, ld = np.linalg.slogdet(full)
    assert abs(s.log_abs_det() - ld) <= LOGDET_RTOL * max(abs(ld), 1.0)
    b = rng.standard_normal(n)
    assert _rel(s.solve(b), np.linalg.solve(full, b)) < SOLVE_RTOL
    assert _rel(s.diag_inv(), np.diag(np.linalg.inv(full))) < SELINV_RTOL


def test_factorize_dense_failure_is_singular_and_keeps_factor():
    # dense_faer.rs:54-67: every dense failure is MathError::Singular; the solver keeps
    # the previous factor (solver.rs:70-74)
    from inla_gpu.solver import Singular

    s = _solver()
    s.factorize_dense(np.eye(3) * 2.0)
    with pytest.raises(Singular):
        s.factorize_dense(np.zeros((3, 3)))
    assert abs(s.log_abs_det() - 3 * np.log(2.0)) < 1e-12
    assert np.allclose(s.solve(np.array([2.0, 4.0, 6.0])), [1.0, 2.0, 3.0])


def test_factorize_dense_empty_matrix():
    # dense_faer.rs:33-39: n = 0 is an empty factor, not an error; the sum of log|D| is 0
    s = _solver()
    s.factorize_dense(np.zeros((0, 0)))
    assert s.n() == 0
    assert s.log_abs_det() == 0.0
    assert s.solve(np.zeros(0)).shape == (0,)
    assert s.diag_inv().shape == (0,)
    # a later factorisation replaces it as usual
    s.factorize_dense(np.eye(2) * 4.0)
    assert abs(s.log_abs_det() - 2 * np.log(4.0)) < 1e-12


def test_sparse_out_of_memory_falls_back_to_dense(monkeypatch):
    # DefaultBackend::factorize (backend.rs:110-122): an out-of-memory sparse
    # factorisation is retried densely. The sparse path's failure is forced from the host
    # (GMRF_SIMULATE_SPARSE_OOM): the factor, a solve and diag(Q^-1) must equal the sparse
    # path's and the oracle's; a failure of the dense retry keeps the previous factor.
    from inla_gpu.solver import Singular

    q = M.random_spd(120, 0.06, 5)
    sp_s = _solver()
    sp_s.factorize(q)
    monkeypatch.setenv("GMRF_SIMULATE_SPARSE_OOM", "1")
    s = _solver()
    s.factorize(q)
    ref = orc.Factor(q)
    b = np.random.default_rng(1).standard_normal(q.shape[0])
    assert abs(s.log_abs_det() - ref.log_abs_det()) <= LOGDET_RTOL * abs(ref.log_abs_det())
    assert abs(s.log_abs_det() - sp_s.log_abs_det()) <= 1e-12 * abs(ref.log_abs_det())
    assert _rel(s.solve(b), ref.solve(b)) < SOLVE_RTOL and _rel(s.solve(b), sp_s.solve(b)) < 1e-12
    assert _rel(s.diag_inv(), ref.takahashi_diag()) < SELINV_RTOL
    zero = sp.csc_matrix(q * 0.0)
    with pytest.raises(Singular):
        s.factorize(zero)
    assert abs(s.log_abs_det() - ref.log_abs_det()) <= LOGDET_RTOL * abs(ref.log_abs_det())
