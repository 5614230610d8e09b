"""gmrf_batch_solve_noalloc (the allocation-free multi-vector solve) against
gmrf_batch_solve_multi through the C ABI: the same kernels on the same inputs, so
the results must be bit-identical; two calls in a row, the second smaller, then a
request beyond the fixed workspace, which must be refused before any work."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_pair(lib, bt, count, nrhs, rhs, n):
    import torch

    x_ref = torch.empty((count * nrhs, n), dtype=torch.float64, device="cuda")
    x_na = torch.full((count * nrhs, n), np.nan, dtype=torch.float64, device="cuda")
    for fn, out in ((lib.gmrf_batch_solve_multi, x_ref), (lib.gmrf_batch_solve_noalloc, x_na)):
        st = fn(bt._h, count, nrhs, C.c_void_p(rhs.data_ptr()), n, C.c_void_p(out.data_ptr()), n)
        assert st == 0, lib.gmrf_last_error()
    assert lib.gmrf_batch_sync(bt._h) == 0
    return x_ref.cpu().numpy(), x_na.cpu().numpy()


def test_noalloc_solve_matches_solve_multi():
    import torch

    from inla_gpu import _lib as L
    from inla_gpu import workloads as W
    from inla_gpu.solver import Batch, Symbolic

    g, _, _ = W.spde_components(30, 30)
    n = g.n
    sym = Symbolic(g.n, g.indptr, g.indices)
    bt = Batch(sym, 3)
    vals = np.stack([g.values(W.spde_coefs(lt, lk)[0]) for lt, lk in ((0.0, 1.5), (0.3, 1.2), (-0.2, 1.8))])
    bt.factorize(torch.as_tensor(vals, device="cuda"))
    _, st = bt.logdet()
    assert (st == 0).all()
    rng = np.random.default_rng(11)
    lib = L.lib
    # first call: 3 matrices x 5 vectors; second call: 2 matrices x 2 vectors
    for count, nrhs in ((3, 5), (2, 2)):
        rhs = torch.as_tensor(rng.standard_normal((count * nrhs, n)), device="cuda")
        ref, got = _solve_pair(lib, bt, count, nrhs, rhs, n)
        print(f"noalloc solve count={count} nrhs={nrhs}: max |diff| = {np.max(np.abs(ref - got)):.3e}")
        assert np.array_equal(ref, got)
        # and they solve the systems
        q0 = g.csc(W.spde_coefs(0.0, 1.5)[0])
        r0 = rhs[0].cpu().numpy()
        assert np.max(np.abs(q0 @ got[0] - r0)) < 1e-9 * np.max(np.abs(r0)) * 1e3
    # beyond the fixed area: refused, nothing written
    per = n + 1  # at least n + 1 doubles per vector
    big = (4 << 20) // per + 1
    nrhs = -(-big // 3)
    assert 3 * nrhs <= 65535
    # buffers sized for the whole request, so that every access would stay in bounds
    rhs = torch.zeros((3 * nrhs, n), dtype=torch.float64, device="cuda")
    out = torch.zeros((3 * nrhs, n), dtype=torch.float64, device="cuda")
    st = lib.gmrf_batch_solve_noalloc(bt._h, 3, nrhs, C.c_void_p(rhs.data_ptr()), 0,
                                      C.c_void_p(out.data_ptr()), 0)
    assert st == L.GMRF_INVALID_ARGUMENT
    st = lib.gmrf_batch_solve_noalloc(bt._h, 3, nrhs, C.c_void_p(rhs.data_ptr()), n,
                                      C.c_void_p(out.data_ptr()), n)
    assert st == L.GMRF_INVALID_ARGUMENT
    assert "allocation-free solve workspace exceeded" in L.last_error()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
