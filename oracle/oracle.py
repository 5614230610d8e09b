"""ORACLE — test infrastructure, not product code.

ctypes wrapper of oracle/_build/liboracle.so (the CPU restatement in
gmrf_oracle.cpp). Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-C", HERE], check=True, capture_output=True)


EVAL_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    i64, dbl, vp, i32 = C.c_int64, C.c_double, C.c_void_p, C.c_int
    p64, pd = C.POINTER(C.c_int64), C.POINTER(C.c_double)
    sig = {
        "orc_lower_bandwidth": (i64, [i64, p64, p64]),
        "orc_bandwidth_with_fwd": (i64, [i64, p64, p64, p64]),
        "orc_choose_order": (i32, [i64, p64, p64, p64]),
        "orc_time_major": (i32, [i64, p64, p64, p64]),
        "orc_rcm": (None, [i64, p64, p64, p64]),
        "orc_min_degree": (None, [i64, p64, p64, p64]),
        "orc_factor": (vp, [i64, p64, p64, pd, p64, C.POINTER(i32)]),
        "orc_free": (None, [vp]),
        "orc_nnz_l": (i64, [vp]),
        "orc_perm": (None, [vp, p64]),
        "orc_d": (None, [vp, pd]),
        "orc_log_abs_det": (dbl, [vp]),
        "orc_solve": (None, [vp, pd, pd]),
        "orc_takahashi_diag": (i32, [vp, pd]),
        "orc_selinv_entries": (i32, [vp, i64, p64, p64, pd]),
        "orc_dense_ldlt": (i32, [i64, pd, pd, pd]),
        "orc_log_gamma": (dbl, [dbl]),
        "orc_eval_lik": (None, [i32, i32, i64, pd, pd, pd, pd, pd, pd, pd]),
        "orc_find_mode": (i32, [i64, p64, p64, pd, i64, p64, p64, pd, i32, i32, pd, pd, pd,
                                i64, pd, pd, i32, dbl, pd, pd, pd, C.POINTER(vp)]),
        "orc_ccd_design": (i32, [i32, dbl, pd, pd]),
        "orc_plant": (None, [i32]),
        "orc_grid_design": (i32, [i32, dbl, dbl, EVAL_FN, vp, pd, i32]),
        "orc_invert_symmetric": (i32, [i32, pd, pd]),
        "orc_jacobi_eigen": (i32, [i32, pd, i32, pd, pd]),
        "orc_normal_cdf": (dbl, [dbl]),
        "orc_poisson_cdf": (dbl, [dbl, i64]),
        "orc_binomial_cdf": (dbl, [i64, dbl, i64]),
        "orc_cpo_failure": (i32, [i32, i32, dbl, dbl, dbl, dbl, dbl]),
        "orc_cpo_pit": (i64, [i32, i32, i32, i64, pd, pd, pd, pd, pd, pd, pd, pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _csc(q):
    q = sp.csc_matrix(q)
    q.sum_duplicates()
    q.sort_indices()
    return (q.shape[0], np.ascontiguousarray(q.indptr, dtype=np.int64),
            np.ascontiguousarray(q.indices, dtype=np.int64), np.ascontiguousarray(q.data, dtype=np.float64))


class OracleError(Exception):
    pass


def lower_bandwidth(q) -> int:
    n, ip, ix, _ = _csc(q)
    return int(lib.orc_lower_bandwidth(n, _p64(ip), _p64(ix)))


def bandwidth_with_fwd(q, fwd) -> int:
    n, ip, ix, _ = _csc(q)
    f = np.ascontiguousarray(fwd, dtype=np.int64)
    return int(lib.orc_bandwidth_with_fwd(n, _p64(ip), _p64(ix), _p64(f)))


def choose_order(q):
    n, ip, ix, _ = _csc(q)
    fwd = np.empty(n, dtype=np.int64)
    k = lib.orc_choose_order(n, _p64(ip), _p64(ix), _p64(fwd))
    return int(k), (fwd if k else None)


def time_major(q):
    n, ip, ix, _ = _csc(q)
    fwd = np.empty(n, dtype=np.int64)
    return fwd if lib.orc_time_major(n, _p64(ip), _p64(ix), _p64(fwd)) else None


def rcm(q):
    n, ip, ix, _ = _csc(q)
    fwd = np.empty(n, dtype=np.int64)
    lib.orc_rcm(n, _p64(ip), _p64(ix), _p64(fwd))
    return fwd


def min_degree(q):
    n, ip, ix, _ = _csc(q)
    out = np.empty(n, dtype=np.int64)
    lib.orc_min_degree(n, _p64(ip), _p64(ix), _p64(out))
    return out


class Factor:
    """Oracle LDL^T of a CSC matrix (perm[k] = original index; None -> chooser / min degree)."""

    def __init__(self, q, perm=None):
        n, ip, ix, x = _csc(q)
        self.n = n
        st = C.c_int(0)
        pp = None
        if perm is not None:
            self._perm = np.ascontiguousarray(perm, dtype=np.int64)
            pp = _p64(self._perm)
        self._h = lib.orc_factor(n, _p64(ip), _p64(ix), _pd(x), pp, C.byref(st))
        self.status = st.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_free(self._h)
            self._h = None

    def log_abs_det(self) -> float:
        return float(lib.orc_log_abs_det(self._h))

    def d(self) -> np.ndarray:
        out = np.empty(self.n)
        lib.orc_d(self._h, _pd(out))
        return out

    def perm(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int64)
        lib.orc_perm(self._h, _p64(out))
        return out

    def nnz_l(self) -> int:
        return int(lib.orc_nnz_l(self._h))

    def solve(self, rhs) -> np.ndarray:
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.empty(self.n)
        lib.orc_solve(self._h, _pd(r), _pd(out))
        return out

    def takahashi_diag(self) -> np.ndarray:
        out = np.empty(self.n)
        st = lib.orc_takahashi_diag(self._h, _pd(out))
        if st:
            raise OracleError(f"takahashi status {st}")
        return out

    def selinv_entries(self, rows, cols) -> np.ndarray:
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        out = np.empty(len(r))
        st = lib.orc_selinv_entries(self._h, len(r), _p64(r), _p64(c), _pd(out))
        if st:
            raise OracleError(f"selinv entry outside the filled pattern (status {st})")
        return out


def dense_ldlt(a: np.ndarray):
    """(log|A|, diag(A^-1)) by the scalar dense LDL^T of ldlt.rs:126-169,232-258."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    ld = C.c_double()
    di = np.empty(n)
    st = lib.orc_dense_ldlt(n, _pd(a), C.byref(ld), _pd(di))
    if st:
        raise OracleError("dense LDL^T singular")
    return ld.value, di


def log_gamma(z: float) -> float:
    return float(lib.orc_log_gamma(z))


FAMILIES = {"gaussian": 0, "poisson": 1, "binomial": 2, "nbinomial": 3}
LINKS = {"identity": 0, "log": 1, "logit": 2}


def eval_lik(family, link, eta, y, a, b=None):
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    m = len(eta)
    y = np.ascontiguousarray(y, dtype=np.float64)
    a = np.ascontiguousarray(np.broadcast_to(a, (m,)), dtype=np.float64)
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, (m,)), dtype=np.float64)
    lp, g, h = np.empty(m), np.empty(m), np.empty(m)
    lib.orc_eval_lik(FAMILIES[family], LINKS[link], m, _pd(eta), _pd(y), _pd(a),
                     None if bb is None else _pd(bb), _pd(lp), _pd(g), _pd(h))
    return lp, g, h


class ModeResult:
    def __init__(self, x, mlik, info, post):
        self.x = x
        self.mlik = mlik
        self.logdet_prior, self.logdet_post, self.n_factorizations, self.n_solves = info
        self._post = post

    def post_factor(self):
        """Oracle handle of the posterior factor (owned here)."""
        f = Factor.__new__(Factor)
        f._h = self._post
        f.n = len(self.x)
        f.status = 0
        self._post = None
        return f

    def __del__(self):
        if getattr(self, "_post", None):
            lib.orc_free(self._post)


def find_mode(q, obs_family, obs_link, y, pa, pb=None, A=None, constraints=None, max_iter=200,
              tol=1e-5, keep_post=False):
    """Laplace mode and mlik at one theta (inference.rs:776-975)."""
    n, ip, ix, x = _csc(q)
    y = np.ascontiguousarray(y, dtype=np.float64)
    nobs = len(y)
    pa = np.ascontiguousarray(np.broadcast_to(pa, (nobs,)), dtype=np.float64)
    pbp = None
    if pb is not None:
        pb = np.ascontiguousarray(np.broadcast_to(pb, (nobs,)), dtype=np.float64)
        pbp = _pd(pb)
    ap = ai = ax = None
    if A is not None:
        A = sp.csc_matrix(A)
        A.sum_duplicates()
        A.sort_indices()
        aip = np.ascontiguousarray(A.indptr, dtype=np.int64)
        aix = np.ascontiguousarray(A.indices, dtype=np.int64)
        axx = np.ascontiguousarray(A.data, dtype=np.float64)
        ap, ai, ax = _p64(aip), _p64(aix), _pd(axx)
    kc = 0
    cap = cep = None
    if constraints is not None:
        ca, ce = constraints
        ca = np.ascontiguousarray(ca, dtype=np.float64).reshape(-1, n)
        ce = np.ascontiguousarray(ce, dtype=np.float64)
        kc = ca.shape[0]
        cap, cep = _pd(ca), _pd(ce)
    xo = np.empty(n)
    ml = C.c_double()
    info = np.empty(4)
    post = C.c_void_p()
    st = lib.orc_find_mode(n, _p64(ip), _p64(ix), _pd(x), nobs, ap, ai, ax, FAMILIES[obs_family],
                           LINKS[obs_link], _pd(y), _pd(pa), pbp, kc, cap, cep, int(max_iter), float(tol),
                           _pd(xo), C.byref(ml), _pd(info), C.byref(post) if keep_post else None)
    if st:
        raise OracleError(f"find_mode status {st}")
    return ModeResult(xo, ml.value, (info[0], info[1], int(info[2]), int(info[3])),
                      post.value if keep_post else None)


def ccd_design(m: int, f0: float = 1.0):
    cap = 1 + 2 * m + 2 ** m
    pts = np.empty(cap * m)
    w = np.empty(cap)
    ne = lib.orc_ccd_design(m, f0, _pd(pts), _pd(w))
    if ne < 0:
        raise OracleError(f"CCD strategy for m={m} is not supported")
    return pts[: ne * m].reshape(ne, m), w[:ne]


# Deliberate mistakes the oracle can switch on (the Plant list of gmrf_oracle.cpp),
# used only by tests/test_oracle_pins.py to show that each pinning check fails when
# the function it pins is wrong.
PLANTS = ["none", "lower_bw", "bw_fwd", "time_major", "rcm", "choose", "min_degree", "ldlt", "logdet",
          "solve", "takahashi", "selinv_lookup", "dense_ldlt", "lgamma", "lik_gauss", "lik_poisson",
          "lik_binomial", "lik_nbinomial", "newton_rhs", "project", "mlik", "kappa", "ccd5", "ccd_weight",
          "grid", "jacobi", "invert", "nnz_l", "perm", "d", "cpo_fail", "cpo", "normal_cdf", "poisson_cdf",
          "binomial_cdf"]


def normal_cdf(x):
    return lib.orc_normal_cdf(float(x))


def poisson_cdf(lam, k):
    return lib.orc_poisson_cdf(float(lam), int(k))


def binomial_cdf(n, p, k):
    return lib.orc_binomial_cdf(int(n), float(p), int(k))


def cpo_failure(family, link, y, a, mu, var, b=0.0):
    """check_cpo_failure_for_obs (model_selection.rs:327-396) for one observation."""
    return bool(lib.orc_cpo_failure(FAMILIES[family], LINKS[link], float(y), float(a), float(b), float(mu),
                                    float(var)))


def cpo_pit(family, link, y, eta, var, weights, a, b=None):
    """compute_cpo_pit (model_selection.rs:222-325): eta, var (K, n); (cpo, pit, failures)
    with NaN where the reference returns None."""
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    K, n = eta.shape
    var = np.ascontiguousarray(np.broadcast_to(var, (K, n)), dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    a = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, (n,)), dtype=np.float64)
    cpo, pit = np.empty(n), np.empty(n)
    nf = lib.orc_cpo_pit(FAMILIES[family], LINKS[link], K, n, _pd(eta), _pd(var), _pd(w), _pd(y), _pd(a),
                         None if bb is None else _pd(bb), _pd(cpo), _pd(pit))
    return cpo, pit, int(nf)


class planted:
    """Context manager: the oracle runs with one planted mistake inside the block."""

    def __init__(self, name):
        self.code = PLANTS.index(name)

    def __enter__(self):
        lib.orc_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib.orc_plant(0)
        return False


def grid_design(m: int, step: float, threshold: float, evaluator):
    """grid_design of integration.rs:64-156 with evaluator(z) -> g(z); (points, weights)."""
    cap = 1 << 16
    pts = np.empty(cap * m)

    def cb(zp, _ctx):
        return float(evaluator(np.array([zp[i] for i in range(m)])))

    fn = EVAL_FN(cb)
    ne = lib.orc_grid_design(m, step, threshold, fn, None, _pd(pts), cap)
    if ne < 0:
        raise OracleError(f"grid design failed ({ne})")
    return pts[: ne * m].reshape(ne, m), np.ones(ne)


def invert_symmetric(h):
    h = np.ascontiguousarray(h, dtype=np.float64)
    m = h.shape[0]
    inv = np.empty((m, m))
    if lib.orc_invert_symmetric(m, _pd(h), _pd(inv)):
        raise OracleError("Singular Hessian matrix")
    return inv


def jacobi_eigen(a, max_iter=100):
    """(eigenvalues, V with eigenvectors as columns) by the Jacobi sweeps of integration.rs:249-318."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m = a.shape[0]
    lam = np.empty(m)
    v = np.empty((m, m))
    if lib.orc_jacobi_eigen(m, _pd(a), int(max_iter), _pd(lam), _pd(v)):
        raise OracleError("Jacobi eigenvalue algorithm did not converge")
    return lam, v
