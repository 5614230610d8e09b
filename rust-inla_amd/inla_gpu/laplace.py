"""Batched Laplace evaluation over theta-points (gmrf_model_* in include/gmrf_c.h).

``LaplaceModel`` holds one latent Gaussian model on the device: an affine
precision Q(theta) = sum_k c_k(theta) Q_k, a projector A, observations and a
likelihood family. ``LaplaceModel.eval`` runs, for a batch of theta-points,
the per-point work of crates/inla_stats/src/inference.rs:776-975 (prior
factorisation, Newton loop, Laplace marginal likelihood), the marginal
variances (inference.rs:1371) and, when coefficient derivatives are given, the
gradient of the marginal likelihood from the selected inverses.

With ``blocks`` the model is a heterogeneous batch of independent models in one
block-diagonal union (gmrf_model_desc.n_blocks): every per-point quantity gains a
block index.

``LogMarginal`` exposes that gradient to torch.autograd.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._lib import EvalArgs, ModelDesc
from .solver import InvalidArgument, as_csc, check

FAMILY = {"gaussian": 0, "poisson": 1, "binomial": 2, "nbinomial": 3}
LINK = {"identity": 0, "log": 1, "logit": 2}

_p64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)


def _arr64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _arrd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dev(t, shape=None):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        a = np.asarray(t, dtype=np.float64)
        t = torch.as_tensor(a if a.flags.writeable else a.copy())
    t = t.to(device="cuda", dtype=torch.float64).contiguous()
    if shape is not None:
        t = t.reshape(shape)
    return t


class LaplaceModel:
    """Latent Gaussian model with affine Q(theta) = sum_k c_k Q_k, resident on the GPU."""

    def __init__(self, n, indptr, indices, components, y, family="gaussian", link=None, A=None,
                 obs_a=None, obs_b=None, constraints=None, capacity=16, constraint_kappa=None,
                 order="chooser", blocks=None, lean=False):
        """constraints = (A_c (k x n), e): hard constraints by projection. With
        constraint_kappa, the augmentation Q + kappa A_c^T A_c of the prior and the
        posterior is applied as an exact low-rank correction (the components then
        must not contain it); otherwise a kappa A_c^T A_c component carries it, as
        the reference forms it. order: "chooser" (the reference's choice), "amd",
        "natural" (the caller's order: a permutation the caller applied beforehand),
        "nd" (nested dissection) or "auto" (the chooser's ordering or nested
        dissection, whichever has the shorter estimated level schedule; same results
        to rounding).
        blocks: nb + 1 latent offsets of independent models in a block-diagonal union
        (observations in block order; see gmrf_model_desc.n_blocks). lean: the prior is
        factorised alone before the Newton loop, so the device holds capacity factors
        instead of 2 capacity (gmrf_model_desc.lean)."""
        link = link or {"gaussian": "identity", "poisson": "log", "binomial": "logit",
                        "nbinomial": "log"}[family]
        self.n = int(n)
        self.indptr = _arr64(indptr)
        self.indices = _arr64(indices)
        comps = _arrd(components)
        if comps.ndim != 2 or comps.shape[1] != len(self.indices):
            raise InvalidArgument("components must be K x nnz")
        self.K = comps.shape[0]
        self.comps = comps
        self.y = _arrd(y)
        self.n_obs = len(self.y)
        self.family, self.link = family, link
        self.capacity = int(capacity)
        keep = [self.indptr, self.indices, comps, self.y]
        d = ModelDesc()
        d.n = self.n
        d.indptr = self.indptr.ctypes.data_as(_p64)
        d.indices = self.indices.ctypes.data_as(_p64)
        d.n_comp = self.K
        d.comps = comps.ctypes.data_as(_pd)
        d.n_obs = self.n_obs
        if A is not None:
            A = as_csc(A)
            if A.shape != (self.n_obs, self.n):
                raise InvalidArgument("A must be n_obs x n")
            self.A = A
            ap, ai, ax = _arr64(A.indptr), _arr64(A.indices), _arrd(A.data)
            keep += [ap, ai, ax]
            d.a_indptr, d.a_indices, d.a_values = ap.ctypes.data_as(_p64), ai.ctypes.data_as(_p64), ax.ctypes.data_as(_pd)
        else:
            self.A = None
        d.family = FAMILY[family]
        d.link = LINK[link]
        d.y = self.y.ctypes.data_as(_pd)
        if obs_a is not None:
            oa = _arrd(np.broadcast_to(obs_a, (self.n_obs,)))
            keep.append(oa)
            d.obs_a = oa.ctypes.data_as(_pd)
        if obs_b is not None:
            ob = _arrd(np.broadcast_to(obs_b, (self.n_obs,)))
            keep.append(ob)
            d.obs_b = ob.ctypes.data_as(_pd)
        if constraints is not None:
            ca, ce = constraints
            ca = _arrd(ca).reshape(-1, self.n)
            ce = _arrd(ce)
            keep += [ca, ce]
            d.n_constr = ca.shape[0]
            d.constr_a = ca.ctypes.data_as(_pd)
            d.constr_e = ce.ctypes.data_as(_pd)
        d.capacity = self.capacity
        if constraint_kappa is not None:
            if constraints is None:
                raise InvalidArgument("constraint_kappa needs constraints")
            d.constr_kappa = float(constraint_kappa)
        d.order = {"chooser": L.ORDER_CHOOSER, "amd": L.ORDER_AMD, "natural": L.ORDER_NATURAL,
                   "nd": L.ORDER_ND, "auto": L.ORDER_AUTO}[order]
        self.nb = 1
        if blocks is not None:
            bp = _arr64(blocks)
            keep.append(bp)
            self.nb = len(bp) - 1
            self.block_ptr = bp
            d.n_blocks = self.nb
            d.block_ptr = bp.ctypes.data_as(_p64)
        d.lean = 1 if lean else 0
        h = C.c_void_p()
        check(L.lib.gmrf_model_create(C.byref(d), C.byref(h)))
        self._h = h
        del keep

    # the free function is bound when the class is defined: at interpreter shutdown
    # the module globals may already be gone when the last handles are collected
    def __del__(self, _free=L.lib.gmrf_model_destroy):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _free(h)

    def eval(self, coefs=None, lpar=None, dcoefs=None, dlog_lpar=None, want_x=True, want_var=True,
             max_iter=200, tol=1e-5, out=None, want_eta=False, qvals=None, dqvals=None):
        """Evaluate a batch of theta-points.

        coefs: (B, K); lpar: (B,) Gaussian precision multipliers; dcoefs: (B, m, K);
        dlog_lpar: (B, m). Returns a dict of device tensors (mlik, x, var, grad) plus
        numpy status / n_iter. ``out``: preallocated float64 CUDA tensors for some of
        mlik (B,), x (B, n), var (B, n), grad (B, m); reusing the same tensors call after
        call lets the library replay its captured graph of the Gaussian evaluation.
        Values mode: qvals (B, nnz) gives Q(theta_b) on the model's pattern instead of
        coefs, and dqvals (B, m, nnz) its theta-derivatives instead of dcoefs
        (gmrf_eval_args.qvals). want_eta adds eta = A x* and eta_var = sum_j A_ij^2 var_j (B, n_obs), the
        predictor means and variances DIC, WAIC and CPO/PIT take (gmrf_eval_args.eta).
        With blocks the per-point arrays carry the block after the point: coefs
        (B, nb, K), lpar (B, nb), dcoefs (B, nb, m, K), dlog_lpar (B, nb, m), mlik
        (B, nb), grad (B, nb, m), status and n_iter (B, nb).
        """
        nb = self.nb
        pp = (nb,) if nb > 1 else ()
        qv = dq = c = dc = None
        m = 0
        if qvals is not None:
            nnz = len(self.indices)
            qv = _dev(qvals)
            if qv.dim() == 1:
                qv = qv.unsqueeze(0)
            B = qv.shape[0]
            if tuple(qv.shape[1:]) != (nnz,):
                raise InvalidArgument("qvals must be B x nnz")
            if dqvals is not None:
                dq = _dev(dqvals).reshape(B, -1, nnz)
                m = dq.shape[1]
        else:
            c = _dev(coefs)
            if c.dim() == 1 + len(pp):
                c = c.unsqueeze(0)
            B = c.shape[0]
            if tuple(c.shape[1:]) != pp + (self.K,):
                raise InvalidArgument("coefs must be B x K" if nb == 1 else "coefs must be B x n_blocks x K")
            dc = _dev(dcoefs) if dcoefs is not None else None
            if dc is not None:
                dc = dc.reshape((B,) + pp + (-1, self.K))
                m = dc.shape[-2]
        if B > self.capacity:
            raise InvalidArgument("batch exceeds model capacity")
        lp = _dev(lpar, (B,) + pp) if lpar is not None else None
        dl = _dev(dlog_lpar, (B,) + pp + (m,)) if dlog_lpar is not None else None
        given = dict(out or {})
        shapes = {"mlik": (B,) + pp, "x": (B, self.n), "var": (B, self.n), "grad": (B,) + pp + (m,),
                  "eta": (B, self.n_obs), "eta_var": (B, self.n_obs)}
        for k, t in given.items():
            if k not in shapes or not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                                       and t.is_contiguous() and tuple(t.shape) == shapes[k]):
                raise InvalidArgument(f"out[{k!r}] must be a contiguous float64 CUDA tensor of shape {shapes[k]}")

        def buf(k):
            return given[k] if k in given else torch.empty(shapes[k], dtype=torch.float64, device="cuda")

        out = {"mlik": buf("mlik")}
        if want_x:
            out["x"] = buf("x")
        if want_var:
            out["var"] = buf("var")
        if dc is not None or dq is not None:
            out["grad"] = buf("grad")
        if want_eta:
            out["eta"] = buf("eta")
            out["eta_var"] = buf("eta_var")
        status = np.zeros((B,) + pp, dtype=np.int32)
        niter = np.zeros((B,) + pp, dtype=np.int32)
        a = EvalArgs()
        a.count, a.n_theta = B, m
        a.coefs = c.data_ptr() if c is not None else None
        a.qvals = qv.data_ptr() if qv is not None else None
        a.dqvals = dq.data_ptr() if dq is not None else None
        a.lpar = lp.data_ptr() if lp is not None else None
        a.dcoefs = dc.data_ptr() if dc is not None else None
        a.dlog_lpar = dl.data_ptr() if dl is not None else None
        a.max_iter, a.tol = int(max_iter), float(tol)
        a.mlik = out["mlik"].data_ptr()
        a.x = out["x"].data_ptr() if want_x else None
        a.var = out["var"].data_ptr() if want_var else None
        a.grad = out["grad"].data_ptr() if "grad" in out else None
        a.status = status.ctypes.data_as(C.POINTER(C.c_int32))
        a.n_iter = niter.ctypes.data_as(C.POINTER(C.c_int32))
        a.eta = out["eta"].data_ptr() if want_eta else None
        a.eta_var = out["eta_var"].data_ptr() if want_eta else None
        check(L.lib.gmrf_model_set_stream(self._h, C.c_void_p(torch.cuda.current_stream().cuda_stream), 1))
        check(L.lib.gmrf_model_eval(self._h, C.byref(a)))
        out["status"] = status
        out["n_iter"] = niter
        return out


    def autotune(self):
        """Tune the launch schedule on the last evaluation's matrices (see
        gmrf_model_autotune); returns (default_ms, tuned_ms) of factorisation +
        selected inversion."""
        d, t = C.c_double(), C.c_double()
        check(L.lib.gmrf_model_autotune(self._h, C.byref(d), C.byref(t)))
        return d.value, t.value


class LogMarginal(torch.autograd.Function):
    """mlik(theta) for a batch of theta-points with the adjoint gradient.

    ``param_fn(theta_np) -> (coefs (B,K), dcoefs (B,m,K), lpar (B,), dlog_lpar (B,m))``
    maps theta to the affine coefficients of Q(theta) and the likelihood
    precision.
    """

    @staticmethod
    def forward(ctx, theta, model, param_fn):
        th = theta.detach().cpu().numpy().reshape(-1, theta.shape[-1])
        coefs, dcoefs, lpar, dlog = param_fn(th)
        r = model.eval(coefs, lpar=lpar, dcoefs=dcoefs, dlog_lpar=dlog, want_x=False, want_var=False)
        ctx.save_for_backward(r["grad"].to(theta.device))
        return r["mlik"].to(theta.device).reshape(theta.shape[:-1])

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout.reshape(-1, 1) * grad, None, None
