"""ctypes binding of libgmrf_hip.so (include/gmrf_c.h).

The shared library is built in-tree next to this file (``__graft_entry__.build``
or ``make -C rust-inla_amd/csrc``). There is no fallback: when the library is
missing, importing this module raises.
"""

from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgmrf_hip.so")

GMRF_OK = 0
GMRF_SINGULAR = 1
GMRF_NOT_POSITIVE_DEFINITE = 2
GMRF_SOLVE_FAILURE = 3
GMRF_INVALID_ARGUMENT = 4
GMRF_OUT_OF_MEMORY = 5
GMRF_DEVICE_ERROR = 6

ORDER_CHOOSER = 0
ORDER_AMD = 1
ORDER_NATURAL = 2
ORDER_GIVEN = 3
ORDER_ND = 4
ORDER_AUTO = 5

_i32 = C.c_int32
_i64 = C.c_int64
_dbl = C.c_double
_vp = C.c_void_p
_pi64 = C.POINTER(C.c_int64)
_pdbl = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)


class SymbolicStats(C.Structure):
    _fields_ = [
        ("n", _i64),
        ("nnz", _i64),
        ("nnz_q_lower", _i64),
        ("order_kind", _i32),
        ("nlevels", _i32),
        ("nsuper", _i64),
        ("nnz_l", _i64),
        ("l_size", _i64),
        ("u_size", _i64),
        ("max_front", _i64),
        ("flops_ldlt", _dbl),
        ("flops_selinv", _dbl),
        ("flops_solve", _dbl),
        ("sum_r2", _dbl),
        ("pattern_hash", C.c_uint64),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("indptr", _pi64),
        ("indices", _pi64),
        ("n_comp", C.c_int32),
        ("comps", _pdbl),
        ("n_obs", C.c_int64),
        ("a_indptr", _pi64),
        ("a_indices", _pi64),
        ("a_values", _pdbl),
        ("family", C.c_int32),
        ("link", C.c_int32),
        ("y", _pdbl),
        ("obs_a", _pdbl),
        ("obs_b", _pdbl),
        ("n_constr", C.c_int32),
        ("constr_a", _pdbl),
        ("constr_e", _pdbl),
        ("capacity", C.c_int32),
        ("constr_kappa", C.c_double),
        ("order", C.c_int32),
        ("n_blocks", C.c_int32),
        ("block_ptr", _pi64),
        ("lean", C.c_int32),
    ]


class EvalArgs(C.Structure):
    _fields_ = [
        ("count", C.c_int32),
        ("n_theta", C.c_int32),
        ("coefs", C.c_void_p),
        ("lpar", C.c_void_p),
        ("dcoefs", C.c_void_p),
        ("dlog_lpar", C.c_void_p),
        ("max_iter", C.c_int32),
        ("tol", C.c_double),
        ("mlik", C.c_void_p),
        ("x", C.c_void_p),
        ("var", C.c_void_p),
        ("grad", C.c_void_p),
        ("status", C.POINTER(C.c_int32)),
        ("n_iter", C.POINTER(C.c_int32)),
        ("eta", C.c_void_p),
        ("eta_var", C.c_void_p),
        ("qvals", C.c_void_p),
        ("dqvals", C.c_void_p),
    ]



# name -> (restype, argtypes); every symbol declared in include/gmrf_c.h
SIGNATURES = {
    "gmrf_last_error": (C.c_char_p, []),
    "gmrf_set_device": (_i32, [_i32]),
    "gmrf_analyze": (_i32, [_i64, _pi64, _pi64, _i32, _pi64, C.POINTER(_vp)]),
    "gmrf_symbolic_free": (None, [_vp]),
    "gmrf_symbolic_get_stats": (_i32, [_vp, C.POINTER(SymbolicStats)]),
    "gmrf_symbolic_perm": (_i32, [_vp, _pi64]),
    "gmrf_symbolic_etree": (_i32, [_vp, _pi64]),
    "gmrf_symbolic_colcount": (_i32, [_vp, _pi64]),
    "gmrf_symbolic_supernodes": (_i32, [_vp, _pi64, _pi64]),
    "gmrf_choose_order": (_i32, [_i64, _pi64, _pi64, _pi64]),
    "gmrf_batch_create": (_i32, [_vp, _i32, C.POINTER(_vp)]),
    "gmrf_batch_destroy": (None, [_vp]),
    "gmrf_batch_factorize": (_i32, [_vp, _i32, _vp, _i64]),
    "gmrf_batch_logdet": (_i32, [_vp, _i32, _pdbl, _pi32]),
    "gmrf_batch_logdet_device": (_i32, [_vp, _i32, _vp, _vp]),
    "gmrf_batch_solve": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64]),
    "gmrf_batch_solve_multi": (_i32, [_vp, _i32, _i32, _vp, _i64, _vp, _i64]),
    "gmrf_batch_solve_noalloc": (_i32, [_vp, _i32, _i32, _vp, _i64, _vp, _i64]),
    "gmrf_batch_autotune": (_i32, [_vp, _i32, _vp, _i64, _pdbl, _pdbl]),
    "gmrf_batch_selinv": (_i32, [_vp, _i32]),
    "gmrf_batch_selinv_diag": (_i32, [_vp, _i32, _vp, _i64]),
    "gmrf_batch_selinv_pattern": (_i32, [_vp, _i32, _vp, _i64]),
    "gmrf_batch_sync": (_i32, [_vp]),
    "gmrf_batch_stream": (_vp, [_vp]),
    "gmrf_batch_set_stream": (_i32, [_vp, _vp, _i32]),
    "gmrf_multi_analyze": (_i32, [_i32, _pi64, C.POINTER(_pi64), C.POINTER(_pi64), _i32, C.POINTER(_vp)]),
    "gmrf_multi_free": (None, [_vp]),
    "gmrf_multi_info": (_i32, [_vp, _pi64, _pi64, _pi32]),
    "gmrf_multi_symbolic": (_vp, [_vp]),
    "gmrf_multi_batch": (_vp, [_vp]),
    "gmrf_multi_factorize": (_i32, [_vp, _i32, _vp, _i64]),
    "gmrf_multi_logdet": (_i32, [_vp, _i32, _vp, _vp]),
    "gmrf_solver_create": (_i32, [C.POINTER(_vp)]),
    "gmrf_solver_destroy": (None, [_vp]),
    "gmrf_solver_factorize": (_i32, [_vp, _i64, _pi64, _pi64, _pdbl]),
    "gmrf_solver_factorize_dense": (_i32, [_vp, _i64, _pdbl]),
    "gmrf_solver_solve": (_i32, [_vp, _pdbl, _pdbl]),
    "gmrf_solver_diag_inv": (_i32, [_vp, _pdbl]),
    "gmrf_solver_log_abs_det": (_i32, [_vp, _pdbl]),
    "gmrf_solver_n": (_i64, [_vp]),
    "gmrf_probe_mma": (_i32, [_pdbl, _pdbl, _i32, _pdbl]),
    "gmrf_model_create": (_i32, [C.POINTER(ModelDesc), C.POINTER(_vp)]),
    "gmrf_model_destroy": (None, [_vp]),
    "gmrf_model_eval": (_i32, [_vp, C.POINTER(EvalArgs)]),
    "gmrf_model_stream": (_vp, [_vp]),
    "gmrf_obs_loglik": (_i32, [_i32, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gmrf_obs_eval": (_i32, [_i32, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gmrf_model_set_stream": (_i32, [_vp, _vp, _i32]),
    "gmrf_model_autotune": (_i32, [_vp, _pdbl, _pdbl]),
    "gmrf_profile_enable": (_i32, [_i32]),
    "gmrf_profile_summary": (_i32, [_i32, C.c_char_p, _i32, _pi64, _pdbl, _pdbl, _pdbl, _dbl, _dbl, _pdbl,
                                    _pi32]),
}


def _load():
    # torch bundles its own HIP runtime with the same SONAME (libamdhip64.so.7).
    # Loading torch first makes this library bind to that runtime instead of
    # pulling a second copy from /opt/rocm into the process.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C rust-inla_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def profile_enable(on: bool) -> None:
    if lib.gmrf_profile_enable(1 if on else 0) != GMRF_OK:
        raise RuntimeError(last_error())


def profile_summary(peak_flops: float = 78.6e12, peak_bytes: float = 8.0e12) -> dict:
    """{kernel: dict(launches, ms, flops, bytes, roof_ms)} of the recorded launches;
    roof_ms = sum over launches of max(flops / peak_flops, bytes / peak_bytes)."""
    cap, nl = 64, 64
    names = C.create_string_buffer(cap * nl)
    la = (C.c_int64 * cap)()
    ms, fl, by, rf = (C.c_double * cap)(), (C.c_double * cap)(), (C.c_double * cap)(), (C.c_double * cap)()
    ne = C.c_int32(0)
    if lib.gmrf_profile_summary(cap, names, nl, la, ms, fl, by, peak_flops, peak_bytes, rf,
                                C.byref(ne)) != GMRF_OK:
        raise RuntimeError(last_error())
    out = {}
    for k in range(min(ne.value, cap)):
        nm = names.raw[k * nl:(k + 1) * nl].split(b"\0", 1)[0].decode()
        out[nm] = {"launches": int(la[k]), "ms": ms[k], "flops": fl[k], "bytes": by[k], "roof_ms": rf[k]}
    return out


def last_error() -> str:
    msg = lib.gmrf_last_error()
    return msg.decode() if msg else ""
