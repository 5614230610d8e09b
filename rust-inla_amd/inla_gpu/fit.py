"""End-to-end fit over theta on the GPU: mode search, Hessian, CCD nodes, mixture.

Two drivers of the same pipeline, both on the batched Laplace evaluator:

* ``fit_batched``: BFGS on the adjoint gradient of the log posterior, with the
  line-search step lengths of an iteration evaluated as one batch; the
  finite-difference Hessian stencil as one batch; all CCD nodes as one batch.
* ``fit_plain``: the reference's sequence one theta-point at a time:
  Nelder-Mead (crates/inla_math/src/hyper_opt.rs:15-185 with step 0.1,
  200 iterations, tol 1e-6 as crates/inla_stats/src/inference.rs:1258), the
  Hessian stencil point by point (hyper_opt.rs:197-259, step from
  inference.rs:1264) and the CCD nodes one by one (inference.rs:1296-1397).

Both end in the mixture of inference.rs:1419-1468. ``strategy`` picks the
integration design as inference.rs:1272-1294 does: "ccd" (ccd_design with f0) or
"grid" (grid_design with step f0 and threshold 4.0, whose evaluator is the log
posterior at the z-points: one batched call per line of the grid in
``fit_batched``, one call per point in ``fit_plain``).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import ccd


@dataclass
class FitResult:
    mode: np.ndarray
    hessian: np.ndarray
    thetas: np.ndarray
    weights: np.ndarray
    log_post: np.ndarray
    mean: np.ndarray
    var: np.ndarray
    mlik: float
    n_evals: int = 0
    n_points: int = 0
    seconds: dict = field(default_factory=dict)


class Objective:
    """log posterior g(theta) = mlik(theta) + log pi(theta) on a LaplaceModel.

    ``params(thetas) -> (coefs, dcoefs, lpar, dlog_lpar)`` and
    ``log_prior(theta) -> (value, gradient)`` describe the model's theta.
    Counts calls (batched calls) and theta-points.
    """

    def __init__(self, model, params, log_prior, max_iter=200, tol=1e-5):
        self.model, self.params, self.log_prior = model, params, log_prior
        self.max_iter, self.tol = max_iter, tol
        self.calls = 0
        self.points = 0

    def _chunks(self, thetas):
        cap = self.model.capacity
        for i in range(0, len(thetas), cap):
            yield thetas[i:i + cap]

    def values(self, thetas, grad=False, want_x=False, want_var=False):
        """g (and its gradient) at each row of thetas, in batches of the model capacity."""
        thetas = np.atleast_2d(np.asarray(thetas, dtype=float))
        out_g, out_d, out_x, out_v = [], [], [], []
        for th in self._chunks(thetas):
            c, dc, lp, dl = self.params(th)
            r = self.model.eval(c, lpar=lp, dcoefs=dc if grad else None, dlog_lpar=dl if grad else None,
                                want_x=want_x, want_var=want_var, max_iter=self.max_iter, tol=self.tol)
            self.calls += 1
            self.points += len(th)
            ml = r["mlik"].cpu().numpy()
            pri = [self.log_prior(t) for t in th]
            out_g.append(ml + np.array([p[0] for p in pri]))
            if grad:
                out_d.append(r["grad"].cpu().numpy() + np.array([p[1] for p in pri]))
            if want_x:
                out_x.append(r["x"].cpu().numpy())
            if want_var:
                out_v.append(r["var"].cpu().numpy())
        res = {"g": np.concatenate(out_g)}
        if grad:
            res["grad"] = np.concatenate(out_d)
        if want_x:
            res["x"] = np.concatenate(out_x)
        if want_var:
            res["var"] = np.concatenate(out_v)
        return res


def _bfgs_steps(theta0, max_iter=100, gtol=1e-3, xtol=1e-6, n_steps=4, max_step=1.0):
    """BFGS maximising g as a generator: it yields arrays of theta-points to evaluate and
    receives (g, grad) at them; its return value is (theta, iterations). Each iteration
    asks for the step lengths 1, 1/2, ..., 2^-(n_steps-1) along the search direction at
    once and takes the longest one that satisfies the Armijo condition. Steps are capped
    at ``max_step`` in every coordinate."""
    x = np.asarray(theta0, dtype=float).copy()
    m = len(x)
    gv0, dv0 = yield x[None]
    f, g = -gv0[0], -dv0[0]  # minimise f = -g
    H = np.eye(m)
    it = 0
    while it < max_iter:
        if np.max(np.abs(g)) < gtol:
            break
        it += 1
        p = -H @ g
        if g @ p >= 0:  # not a descent direction: restart from steepest descent
            H = np.eye(m)
            p = -g
        big = np.max(np.abs(p))
        if big > max_step:  # theta lives on a log / logit scale
            p = p * (max_step / big)
        ts = 0.5 ** np.arange(n_steps)
        accepted = None
        while accepted is None:
            gv, dv = yield x[None, :] + ts[:, None] * p[None, :]
            fv, gvv = -np.asarray(gv), -np.asarray(dv)
            for k, t in enumerate(ts):
                if np.isfinite(fv[k]) and fv[k] <= f + 1e-4 * t * (g @ p):
                    accepted = (t, fv[k], gvv[k])
                    break
            if accepted is None:
                ts = ts * 0.5 ** n_steps
                if ts[0] < 1e-12:
                    return x, it
        t, fn, gn = accepted
        s = t * p
        yk = gn - g
        sy = s @ yk
        if sy > 1e-12 * max(1.0, np.linalg.norm(s) * np.linalg.norm(yk)):
            if it == 1:
                H = np.eye(m) * (sy / (yk @ yk))
            rho = 1.0 / sy
            I = np.eye(m)
            H = (I - rho * np.outer(s, yk)) @ H @ (I - rho * np.outer(yk, s)) + rho * np.outer(s, s)
        x, f, g = x + s, fn, gn
        if np.max(np.abs(s)) < xtol:
            break
    return x, it


def bfgs_mode(obj: Objective, theta0, **kw):
    """Maximise g with BFGS on the adjoint gradient (_bfgs_steps), each request one
    batched evaluation. Returns (theta, iterations)."""
    gen = _bfgs_steps(theta0, **kw)
    pts = next(gen)
    while True:
        r = obj.values(pts, grad=True)
        try:
            pts = gen.send((r["g"], r["grad"]))
        except StopIteration as e:
            return e.value


def nelder_mead_mode(obj: Objective, theta0, step=0.1, max_iter=200, tol=1e-6):
    """Minimise -g with the downhill simplex of hyper_opt.rs:15-185, one point per call."""
    def f(th):
        v = -obj.values(np.asarray(th)[None])["g"][0]
        if not np.isfinite(v):
            raise RuntimeError("objective is not finite")
        return v

    m = len(theta0)
    verts = [np.asarray(theta0, dtype=float).copy()]
    fv = [f(verts[0])]
    for i in range(m):
        v = verts[0].copy()
        v[i] += step
        verts.append(v)
        fv.append(f(v))
    for _ in range(max_iter):
        order = sorted(range(m + 1), key=lambda i: fv[i])
        verts = [verts[i] for i in order]
        fv = [fv[i] for i in order]
        if np.std(fv) < tol:  # population std, as the reference
            return verts[0]
        cen = np.mean(verts[:m], axis=0)
        refl = cen + (cen - verts[m])
        fr = f(refl)
        if fv[0] <= fr < fv[m - 1]:
            verts[m], fv[m] = refl, fr
            continue
        if fr < fv[0]:
            exp_ = cen + 2.0 * (refl - cen)
            fe = f(exp_)
            if fe < fr:
                verts[m], fv[m] = exp_, fe
            else:
                verts[m], fv[m] = refl, fr
            continue
        if fr < fv[m]:
            con = cen + 0.5 * (refl - cen)
            fc = f(con)
            if fc < fr:
                verts[m], fv[m] = con, fc
                continue
        else:
            con = cen + 0.5 * (verts[m] - cen)
            fc = f(con)
            if fc < fv[m]:
                verts[m], fv[m] = con, fc
                continue
        for i in range(1, m + 1):
            verts[i] = verts[0] + 0.5 * (verts[i] - verts[0])
            fv[i] = f(verts[i])
    return verts[0]


def _hessian_step(mode):
    return max(1e-4, 0.05 * float(np.max(np.abs(mode))))


GRID_THRESHOLD = 4.0  # inference.rs:1291


def _nodes(mode, H, strategy="ccd", f0=1.0, evaluate=None):
    """theta-points and design weights around the mode; ``evaluate(thetas) -> g`` is
    the log posterior the grid design walks."""
    lam, V = ccd.scaling_from_hessian(H)

    def to_theta(Z):
        return np.array([ccd.theta_from_z(mode, lam, V, zz) for zz in np.atleast_2d(Z)])

    if strategy == "grid":
        z, w = ccd.grid_design(len(mode), f0, GRID_THRESHOLD, lambda Z: evaluate(to_theta(Z)))
    elif strategy == "ccd":
        z, w = ccd.ccd_design(len(mode), f0)
    else:
        raise ValueError(f"unknown integration strategy {strategy!r}")
    return to_theta(z), w


def _finish(mode, H, thetas, w, r, n_evals, n_points, secs):
    wts, mean, var, mlik = ccd.aggregate(w, r["g"], r["x"], r["var"])
    return FitResult(mode, H, thetas, wts, r["g"], mean, var, mlik, n_evals, n_points, secs)


def fit_batched(obj: Objective, theta0, strategy="ccd", f0=1.0, **bfgs_kw) -> FitResult:
    """Mode by batched BFGS, Hessian stencil and the design's nodes as single batches."""
    secs = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mode, _ = bfgs_mode(obj, theta0, **bfgs_kw)
    t1 = time.perf_counter()
    h = _hessian_step(mode)
    pts = ccd.hessian_points(mode, h)
    H = ccd.hessian_from_values(obj.values(pts)["g"], len(mode), h)
    t2 = time.perf_counter()
    thetas, w = _nodes(mode, H, strategy, f0, lambda th: obj.values(th)["g"])
    r = obj.values(thetas, want_x=True, want_var=True)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    secs.update(mode=t1 - t0, hessian=t2 - t1, nodes=t3 - t2, total=t3 - t0)
    return _finish(mode, H, thetas, w, r, obj.calls, obj.points, secs)


def fit_plain(obj: Objective, theta0, strategy="ccd", f0=1.0) -> FitResult:
    """The reference's sequence, one theta-point per evaluation."""
    secs = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mode = nelder_mead_mode(obj, theta0)
    t1 = time.perf_counter()
    h = _hessian_step(mode)
    pts = ccd.hessian_points(mode, h)
    H = ccd.hessian_from_values(np.concatenate([obj.values(p[None])["g"] for p in pts]), len(mode), h)
    t2 = time.perf_counter()
    thetas, w = _nodes(mode, H, strategy, f0,
                       lambda th: np.concatenate([obj.values(t[None])["g"] for t in th]))
    parts = [obj.values(t[None], want_x=True, want_var=True) for t in thetas]
    r = {k: np.concatenate([p[k] for p in parts]) for k in ("g", "x", "var")}
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    secs.update(mode=t1 - t0, hessian=t2 - t1, nodes=t3 - t2, total=t3 - t0)
    return _finish(mode, H, thetas, w, r, obj.calls, obj.points, secs)


# ---------------------------------------------------------------------------
# Independent models as the blocks of one LaplaceModel
# ---------------------------------------------------------------------------
class BlockObjective:
    """log posteriors g_q(theta_q) = mlik_q + log pi_q(theta_q) of k independent models,
    the blocks of one LaplaceModel (blocks=...): one evaluation takes theta-points of
    shape (B, k, m), model q at its own theta[b, q].

    ``params(thetas (B, k, m)) -> (coefs (B,k,K), dcoefs (B,k,m,K), lpar, dlog_lpar)``;
    ``log_prior(theta (m,)) -> (value, gradient)`` (the same prior for every model)."""

    def __init__(self, model, params, log_prior, max_iter=200, tol=1e-5):
        self.model, self.params, self.log_prior = model, params, log_prior
        self.k = model.nb
        self.max_iter, self.tol = max_iter, tol
        self.calls = 0
        self.points = 0

    def values(self, thetas, grad=False, want_x=False, want_var=False):
        thetas = np.asarray(thetas, dtype=float)
        if thetas.ndim == 2:
            thetas = thetas[None]
        acc = {k: [] for k in ("g", "grad", "x", "var", "status")}
        cap = self.model.capacity
        for i in range(0, len(thetas), cap):
            th = thetas[i:i + cap]
            c, dc, lp, dl = self.params(th)
            r = self.model.eval(c, lpar=lp, dcoefs=dc if grad else None, dlog_lpar=dl if grad else None,
                                want_x=want_x, want_var=want_var, max_iter=self.max_iter, tol=self.tol)
            self.calls += 1
            self.points += th.shape[0] * th.shape[1]
            pri = [[self.log_prior(t) for t in row] for row in th]
            acc["g"].append(r["mlik"].cpu().numpy() + np.array([[p[0] for p in row] for row in pri]))
            acc["status"].append(r["status"])
            if grad:
                acc["grad"].append(r["grad"].cpu().numpy() + np.array([[p[1] for p in row] for row in pri]))
            if want_x:
                acc["x"].append(r["x"].cpu().numpy())
            if want_var:
                acc["var"].append(r["var"].cpu().numpy())
        return {k: np.concatenate(v) for k, v in acc.items() if v}


def bfgs_blocks(obj: BlockObjective, theta0, **kw):
    """One BFGS (_bfgs_steps) per model, all advanced together: each round gathers the
    theta-points every running search asks for into one (B, k, m) evaluation (padded
    with the model's last point; finished models repeat their mode). Returns (modes
    (k, m), iterations (k,), rounds)."""
    theta0 = np.asarray(theta0, dtype=float)
    k = theta0.shape[0]
    gens = [_bfgs_steps(theta0[q], **kw) for q in range(k)]
    want = [next(g) for g in gens]
    done = [None] * k
    rounds = 0
    while any(d is None for d in done):
        rounds += 1
        nb = max(len(w) for q, w in enumerate(want) if done[q] is None)
        T = np.empty((nb, k, theta0.shape[1]))
        for q in range(k):
            src = want[q] if done[q] is None else done[q][0][None]
            T[:, q] = np.concatenate([src, np.repeat(src[-1:], nb - len(src), axis=0)])
        r = obj.values(T, grad=True)
        for q in range(k):
            if done[q] is not None:
                continue
            nq = len(want[q])
            try:
                want[q] = gens[q].send((r["g"][:nq, q], r["grad"][:nq, q]))
            except StopIteration as e:
                done[q] = e.value
    return np.array([d[0] for d in done]), np.array([d[1] for d in done]), rounds


def fit_blocks(obj: BlockObjective, theta0, block_ptr, f0=1.0, **bfgs_kw):
    """Every model's fit at once: BFGS modes (bfgs_blocks), each model's Hessian stencil
    in one evaluation, each model's CCD nodes in one evaluation, and the mixture per
    model (inference.rs:1246-1468 per model). Returns a list of FitResult, one per
    model (latent moments over the model's own latent segment), and the timings."""
    secs = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    modes, _, rounds = bfgs_blocks(obj, theta0, **bfgs_kw)
    t1 = time.perf_counter()
    k, m = modes.shape
    hs = [_hessian_step(md) for md in modes]
    st = np.stack([np.asarray(ccd.hessian_points(modes[q], hs[q])) for q in range(k)], axis=1)
    gs = obj.values(st)["g"]
    Hs = [ccd.hessian_from_values(gs[:, q], m, hs[q]) for q in range(k)]
    t2 = time.perf_counter()
    nodes = [_nodes(modes[q], Hs[q], "ccd", f0) for q in range(k)]
    T = np.stack([nd[0] for nd in nodes], axis=1)
    r = obj.values(T, want_x=True, want_var=True)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = []
    for q in range(k):
        a, b = block_ptr[q], block_ptr[q + 1]
        rq = {"g": r["g"][:, q], "x": r["x"][:, a:b], "var": r["var"][:, a:b]}
        out.append(_finish(modes[q], Hs[q], T[:, q], nodes[q][1], rq, obj.calls, obj.points, {}))
    secs.update(mode=t1 - t0, hessian=t2 - t1, nodes=t3 - t2, total=t3 - t0, bfgs_rounds=rounds)
    return out, secs
