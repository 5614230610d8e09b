"""Python mirror of rust-inla's solver interface on the HIP library.

``HipSolver`` follows ``trait InlaSolver`` (crates/inla_math/src/solver.rs:29-44)
method for method — ``factorize``, ``solve``, ``diag_inv``, ``log_abs_det`` —
with the same error variants (``SolverError``: Singular, NotPositiveDefinite,
SolveFailure; solver.rs:10-17). ``Symbolic`` and ``Batch`` expose the batched
numeric phase over one symbolic analysis on device tensors (torch, float64).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L


class SolverError(Exception):
    """Base of the solver errors (SolverError in solver.rs:10-17)."""


class Singular(SolverError):
    """Matrix is singular or structurally rank-deficient."""


class NotPositiveDefinite(SolverError):
    """Matrix is not positive-definite during factorization."""


class SolveFailure(SolverError):
    """Linear system solve failed."""


class InvalidArgument(SolverError, ValueError):
    """Shape / pattern errors (MathError::NotSquare, DimensionMismatch, ...)."""


class OutOfMemory(SolverError, MemoryError):
    """Device allocation failed."""


class DeviceError(SolverError, RuntimeError):
    """HIP runtime failure."""


_ERRORS = {
    L.GMRF_SINGULAR: Singular,
    L.GMRF_NOT_POSITIVE_DEFINITE: NotPositiveDefinite,
    L.GMRF_SOLVE_FAILURE: SolveFailure,
    L.GMRF_INVALID_ARGUMENT: InvalidArgument,
    L.GMRF_OUT_OF_MEMORY: OutOfMemory,
    L.GMRF_DEVICE_ERROR: DeviceError,
}


def check(status: int) -> None:
    if status != L.GMRF_OK:
        raise _ERRORS.get(status, SolverError)(L.last_error())


def _i64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def as_csc(q) -> sp.csc_matrix:
    """Canonical CSC (sorted, duplicates summed, int64 indices)."""
    q = sp.csc_matrix(q)
    if not q.has_canonical_format:
        q = q.copy()
        q.sum_duplicates()
        q.sort_indices()
    return q


class Symbolic:
    """Symbolic analysis of one CSC pattern (ordering, etree, supernodes)."""

    def __init__(self, n: int, indptr, indices, order: int = L.ORDER_CHOOSER, perm=None):
        self.n = int(n)
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        p = None
        if perm is not None:
            self._perm_in = np.ascontiguousarray(perm, dtype=np.int64)
            p = _i64p(self._perm_in)
        h = C.c_void_p()
        check(L.lib.gmrf_analyze(self.n, _i64p(self.indptr), _i64p(self.indices), int(order), p,
                                 C.byref(h)))
        self._h = h

    @classmethod
    def from_csc(cls, q, **kw):
        q = as_csc(q)
        return cls(q.shape[0], q.indptr, q.indices, **kw)

    # the free function is bound when the class is defined: at interpreter shutdown
    # the module globals may already be gone when the last handles are collected
    def __del__(self, _free=L.lib.gmrf_symbolic_free):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _free(h)

    @property
    def handle(self):
        return self._h

    def stats(self) -> dict:
        st = L.SymbolicStats()
        check(L.lib.gmrf_symbolic_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def perm(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int64)
        check(L.lib.gmrf_symbolic_perm(self._h, _i64p(out)))
        return out

    def etree(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int64)
        check(L.lib.gmrf_symbolic_etree(self._h, _i64p(out)))
        return out

    def colcount(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int64)
        check(L.lib.gmrf_symbolic_colcount(self._h, _i64p(out)))
        return out

    def supernodes(self):
        ns = self.stats()["nsuper"]
        start = np.empty(ns + 1, dtype=np.int64)
        par = np.empty(ns, dtype=np.int64)
        check(L.lib.gmrf_symbolic_supernodes(self._h, _i64p(start), _i64p(par)))
        return start, par


def choose_order(q):
    """Restated choose_symmetric_order: (branch, fwd or None); branch 0 AMD, 1 time-major, 2 RCM."""
    q = as_csc(q)
    n = q.shape[0]
    ip = np.ascontiguousarray(q.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(q.indices, dtype=np.int64)
    fwd = np.empty(n, dtype=np.int64)
    k = L.lib.gmrf_choose_order(n, _i64p(ip), _i64p(ix), _i64p(fwd))
    return int(k), (fwd if k else None)


class Batch:
    """Up to ``capacity`` numeric factorisations sharing one ``Symbolic``.

    All array arguments are contiguous float64 torch tensors on the library's
    device; work is enqueued on torch's current stream.
    """

    def __init__(self, sym: Symbolic, capacity: int):
        self.sym = sym
        self.capacity = int(capacity)
        self.n = sym.n
        self.nnz = int(sym.indptr[-1])
        h = C.c_void_p()
        check(L.lib.gmrf_batch_create(sym.handle, self.capacity, C.byref(h)))
        self._h = h

    # the free function is bound when the class is defined: at interpreter shutdown
    # the module globals may already be gone when the last handles are collected
    def __del__(self, _free=L.lib.gmrf_batch_destroy):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _free(h)

    def _stream(self):
        import torch

        check(L.lib.gmrf_batch_set_stream(self._h, C.c_void_p(torch.cuda.current_stream().cuda_stream), 1))

    @staticmethod
    def _dev(t, cols, name):
        import torch

        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
            raise InvalidArgument(f"{name} must be a float64 CUDA tensor")
        if t.dim() == 1:
            t = t.unsqueeze(0)
        if t.shape[-1] != cols or not t.is_contiguous():
            raise InvalidArgument(f"{name} must be contiguous with {cols} columns")
        return t

    def factorize(self, values):
        v = self._dev(values, self.nnz, "values")
        self._stream()
        check(L.lib.gmrf_batch_factorize(self._h, v.shape[0], C.c_void_p(v.data_ptr()), self.nnz))
        self.count = v.shape[0]
        return self

    def logdet(self, count=None):
        """(log|Q_b| as numpy, status codes as numpy); synchronises."""
        import torch

        count = self.count if count is None else count
        torch.cuda.current_stream().synchronize()
        ld = np.empty(count)
        st = np.empty(count, dtype=np.int32)
        check(L.lib.gmrf_batch_logdet(self._h, count, _dp(ld), st.ctypes.data_as(C.POINTER(C.c_int32))))
        return ld, st

    def logdet_device(self, out_logdet, out_status):
        self._stream()
        check(L.lib.gmrf_batch_logdet_device(self._h, out_logdet.shape[0], C.c_void_p(out_logdet.data_ptr()),
                                             C.c_void_p(out_status.data_ptr())))

    def solve(self, rhs, out=None):
        import torch

        r = self._dev(rhs, self.n, "rhs")
        if out is None:
            out = torch.empty_like(r)
        self._stream()
        check(L.lib.gmrf_batch_solve(self._h, r.shape[0], C.c_void_p(r.data_ptr()), self.n,
                                     C.c_void_p(out.data_ptr()), self.n))
        return out

    def solve_multi(self, rhs, out=None):
        """rhs of shape (count, nrhs, n): nrhs right-hand sides per matrix, one launch sequence."""
        import torch

        if rhs.dim() != 3 or rhs.shape[2] != self.n:
            raise ValueError("rhs must have shape (count, nrhs, n)")
        r = self._dev(rhs.reshape(-1, self.n), self.n, "rhs").reshape(rhs.shape)
        if out is None:
            out = torch.empty_like(r)
        self._stream()
        check(L.lib.gmrf_batch_solve_multi(self._h, r.shape[0], r.shape[1], C.c_void_p(r.data_ptr()), self.n,
                                           C.c_void_p(out.data_ptr()), self.n))
        return out

    def autotune(self, values):
        """Tune the launch schedule for these matrices (gmrf_batch_autotune); the
        factors are overwritten. Returns (default_ms, tuned_ms)."""
        v = self._dev(values, self.nnz, "values")
        self._stream()
        d, t = C.c_double(), C.c_double()
        check(L.lib.gmrf_batch_autotune(self._h, v.shape[0], C.c_void_p(v.data_ptr()), self.nnz, C.byref(d),
                                        C.byref(t)))
        return d.value, t.value

    def selinv(self, count=None):
        count = self.count if count is None else count
        self._stream()
        check(L.lib.gmrf_batch_selinv(self._h, count))
        return self

    def selinv_diag(self, count=None):
        import torch

        count = self.count if count is None else count
        out = torch.empty((count, self.n), dtype=torch.float64, device="cuda")
        self._stream()
        check(L.lib.gmrf_batch_selinv_diag(self._h, count, C.c_void_p(out.data_ptr()), self.n))
        return out

    def selinv_pattern(self, count=None):
        import torch

        count = self.count if count is None else count
        out = torch.empty((count, self.nnz), dtype=torch.float64, device="cuda")
        self._stream()
        check(L.lib.gmrf_batch_selinv_pattern(self._h, count, C.c_void_p(out.data_ptr()), self.nnz))
        return out


class MultiBatch(Batch):
    """Heterogeneous batch: k matrices with different patterns factorised as
    their block-diagonal union in one launch sequence (gmrf_multi_* in
    include/gmrf_c.h). Values, right-hand sides and outputs use the union
    layout (the per-pattern arrays concatenated in order); ``logdet`` returns
    one log-determinant per pattern."""

    def __init__(self, mats, capacity: int = 1):
        mats = [as_csc(q) for q in mats]
        self.k = len(mats)
        if self.k < 1:
            raise InvalidArgument("need at least one pattern")
        self._keep = [(np.ascontiguousarray(q.indptr, dtype=np.int64),
                       np.ascontiguousarray(q.indices, dtype=np.int64)) for q in mats]
        ns = np.array([q.shape[0] for q in mats], dtype=np.int64)
        pi = C.POINTER(C.c_int64)
        ip = (pi * self.k)(*[_i64p(a) for a, _ in self._keep])
        ix = (pi * self.k)(*[_i64p(b) for _, b in self._keep])
        h = C.c_void_p()
        check(L.lib.gmrf_multi_analyze(self.k, _i64p(ns), ip, ix, int(capacity), C.byref(h)))
        self._m = h
        self.n_off = np.empty(self.k + 1, dtype=np.int64)
        self.nnz_off = np.empty(self.k + 1, dtype=np.int64)
        self.order_kind = np.empty(self.k, dtype=np.int32)
        check(L.lib.gmrf_multi_info(h, _i64p(self.n_off), _i64p(self.nnz_off),
                                    self.order_kind.ctypes.data_as(C.POINTER(C.c_int32))))
        self.capacity = int(capacity)
        self.n = int(self.n_off[-1])
        self.nnz = int(self.nnz_off[-1])
        self.sym = None
        self._h = C.c_void_p(L.lib.gmrf_multi_batch(h))  # owned by the multi handle

    def __del__(self, _free=L.lib.gmrf_multi_free):
        m = getattr(self, "_m", None)
        if m:
            self._h = None
            self._m = None
            _free(m)

    def factorize(self, values):
        v = self._dev(values, self.nnz, "values")
        self._stream()
        check(L.lib.gmrf_multi_factorize(self._m, v.shape[0], C.c_void_p(v.data_ptr()), self.nnz))
        self.count = v.shape[0]
        return self

    def logdet(self, count=None):
        """(log|Q_q| per matrix and pattern, status), numpy arrays of shape (count, k)."""
        import torch

        count = self.count if count is None else count
        torch.cuda.current_stream().synchronize()
        ld = np.empty((count, self.k))
        st = np.empty((count, self.k), dtype=np.int32)
        check(L.lib.gmrf_multi_logdet(self._m, count, C.c_void_p(ld.ctypes.data), C.c_void_p(st.ctypes.data)))
        return ld, st

    def split(self, t):
        """Per-pattern views of a union-layout vector tensor (last axis n)."""
        return [t[..., self.n_off[q]:self.n_off[q + 1]] for q in range(self.k)]


class HipSolver:
    """InlaSolver on the GPU (drop-in for FaerCpuSolver, solver.rs:48-99)."""

    def __init__(self):
        h = C.c_void_p()
        check(L.lib.gmrf_solver_create(C.byref(h)))
        self._h = h

    # the free function is bound when the class is defined: at interpreter shutdown
    # the module globals may already be gone when the last handles are collected
    def __del__(self, _free=L.lib.gmrf_solver_destroy):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _free(h)

    def factorize(self, q) -> None:
        q = as_csc(q)
        n, m = q.shape
        if n != m:
            raise InvalidArgument(f"matrix must be square (got {n}x{m})")
        ip = np.ascontiguousarray(q.indptr, dtype=np.int64)
        ix = np.ascontiguousarray(q.indices, dtype=np.int64)
        x = np.ascontiguousarray(q.data, dtype=np.float64)
        check(L.lib.gmrf_solver_factorize(self._h, n, _i64p(ip), _i64p(ix), _dp(x)))

    def factorize_dense(self, a) -> None:
        """LdltBackend::factorize_dense (backend.rs:59-69): a is n x n, lower triangle read."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise InvalidArgument(f"dimension mismatch: dense LDLᵀ matrix must be square (got {a.shape})")
        check(L.lib.gmrf_solver_factorize_dense(self._h, a.shape[0], _dp(a)))

    def n(self) -> int:
        return int(L.lib.gmrf_solver_n(self._h))

    def solve(self, rhs) -> np.ndarray:
        n = self.n()
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        if n and r.shape != (n,):
            raise InvalidArgument(
                f"dimension mismatch: sparse LDLᵀ solve RHS (expected {n}, got {r.shape[0]})")
        out = np.empty_like(r)
        check(L.lib.gmrf_solver_solve(self._h, _dp(r), _dp(out)))
        return out

    def diag_inv(self) -> np.ndarray:
        out = np.empty(max(self.n(), 1))
        check(L.lib.gmrf_solver_diag_inv(self._h, _dp(out)))
        return out[: self.n()]

    def log_abs_det(self) -> float:
        v = C.c_double()
        check(L.lib.gmrf_solver_log_abs_det(self._h, C.byref(v)))
        return v.value


def probe_mma(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """T = A B^T through the MFMA tile engine (A, B: 64 x K, K <= 64)."""
    K = A.shape[1]
    a = np.asfortranarray(A, dtype=np.float64)
    b = np.asfortranarray(B, dtype=np.float64)
    t = np.empty((64, 64), dtype=np.float64, order="F")
    check(L.lib.gmrf_probe_mma(_dp(a), _dp(b), K, _dp(t)))
    return t
