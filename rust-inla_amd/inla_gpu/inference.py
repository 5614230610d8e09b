"""run_inla_inference: py-inla's entry point (py-inla/src/lib.rs:427-657, over
crates/inla_stats/src/inference.rs:1246-1597) on the batched GPU evaluation.

The arguments and the result follow the reference's contract:
- ``build_prior(theta) -> CSC`` and ``log_prior_density(theta) -> float`` callbacks;
- observation dicts;
- the CCD or grid strategy;
- optional A, constraints, Gaussian free precision, DIC / WAIC / CPO;
- latent and predictor marginals.

The work behind them is this package's:
- Q(theta) goes to the device as values on one fixed pattern (values mode of
  gmrf_model_eval);
- every batch of theta-points (the simplex's trial points one by one, as the reference
  does, or a BFGS line search with ``mode_search="bfgs"``; the Hessian stencil; the
  design's nodes) is one batched Laplace evaluation;
- the gradient for BFGS is the adjoint one, with dQ/dtheta from central differences of
  ``build_prior``.

Single-family observation lists are supported: Gaussian, Poisson, binomial and negative
binomial, each with the reference's default link or an explicit one.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import ccd
from . import fit as F
from . import model_selection as MS
from .laplace import LaplaceModel
from .solver import InvalidArgument, as_csc

HARD_CONSTRAINT_KAPPA = 3.059023205018258e6  # constraints.rs:268
_DEFAULT_LINK = {"gaussian": "identity", "poisson": "log", "binomial": "logit", "nbinomial": "log"}
_N_SD = 4.0  # MarginalOptions::default (marginals.rs)


@dataclass
class Marginal1D:
    x: np.ndarray
    y: np.ndarray


@dataclass
class InferenceResult:
    """The fields of the reference's InferenceResult (inference.rs:978-1024)."""

    mode: np.ndarray
    hessian: np.ndarray
    latent_means: np.ndarray
    latent_variances: np.ndarray
    predictor_means: np.ndarray
    predictor_variances: np.ndarray
    marginal_log_lik: float
    marginal_log_lik_gaussian: float
    dic: float
    mean_deviance: float
    effective_params: float
    waic: float
    waic_lppd: float
    waic_effective_params: float
    cpo: list
    pit: list
    cpo_n_failures: int
    theta_nodes: np.ndarray
    node_weights: np.ndarray
    internal_marginals_hyperpar: list
    marginals_latent: list
    marginals_latent_indices: list
    marginals_predictor: list
    marginals_predictor_indices: list
    posterior_precision: sp.csc_matrix | None
    log_posts: np.ndarray = None
    evaluations: int = 0
    extra: dict = field(default_factory=dict)

    def lincomb(self, combs):
        """Gaussian summaries of v = a^T x (the binding's lincomb, py-inla/src/lib.rs:339-366):
        combs is ``[(name, [(latent_index, weight), ...]), ...]``; the variance comes from
        the posterior precision stored at the mode (posterior.rs:53-91)."""
        if self.posterior_precision is None:
            raise InvalidArgument("lincomb: posterior precision was not stored")
        from .posterior import lincomb_summaries

        return lincomb_summaries(self.latent_means, self.posterior_precision, combs)


# --------------------------------------------------------------------------- observations
def parse_obs(obs, gaussian_free_prec=False):
    """The observation dicts of parse_obs (py-inla/src/lib.rs:1115-1260) for one family:
    (family, link, y, obs_a, obs_b, prec_slot). obs_a is the Gaussian precision, the
    Poisson exposure ("exposure" or "E"), the binomial trials ("n") or the NB exposure;
    obs_b the NB size. With gaussian_free_prec the Gaussian precision is exp(theta[slot])
    with slot = "prec_theta" (default 0) for every observation, as the reference's
    build_obs closure."""
    if len(obs) == 0:
        raise InvalidArgument("no observations")
    fams, links = set(), set()
    y, a, b, slots = [], [], [], set()
    for o in obs:
        if o is None or o.get("family") is None or o.get("family") == "none":
            raise InvalidArgument("observations without a family (Obs::None) are not supported")
        fam = {"negative_binomial": "nbinomial"}.get(o["family"], o["family"])
        if fam not in _DEFAULT_LINK:
            raise InvalidArgument(f"unsupported observation family: {o['family']}")
        link = o.get("link") or "default"
        link = _DEFAULT_LINK[fam] if link in ("default", "") else link
        if link not in ("identity", "log", "logit"):
            raise InvalidArgument(f"unknown link function: {link}")
        fams.add(fam)
        links.add(link)
        y.append(float(o["y"]))
        if fam == "gaussian":
            a.append(float(o["precision"]))
            slots.add(o.get("prec_theta"))
        elif fam == "poisson":
            e = o.get("exposure")
            a.append(float(o["E"] if e is None else e))
        elif fam == "binomial":
            a.append(float(o["n"]))
        else:
            a.append(float(o["exposure"]))
            b.append(float(o["size"]))
    if len(fams) != 1 or len(links) != 1:
        raise InvalidArgument("mixed observation families or links are not supported")
    fam, link = fams.pop(), links.pop()
    slot = None
    if fam == "gaussian" and gaussian_free_prec:
        if len(slots) != 1:
            raise InvalidArgument("gaussian_free_prec needs one prec_theta slot for all observations")
        s = slots.pop()
        slot = 0 if s is None else int(s)
    return fam, link, np.array(y), np.array(a), (np.array(b) if b else None), slot


# --------------------------------------------------------------------------- node evaluator
class GpuNodes:
    """Batched Laplace node evaluations for callbacks: the Objective interface of
    inla_gpu.fit (``values(thetas, grad, want_x, want_var)`` -> log posterior g and its
    gradient) plus the predictor outputs, in values mode on one pattern."""

    def __init__(self, build_prior, log_prior_density, m, family, link, y, obs_a, obs_b, slot, A=None,
                 constraints=None, capacity=16, theta0=None):
        self.build_prior, self.log_prior_density = build_prior, log_prior_density
        self.m, self.family, self.link, self.slot = m, family, link, slot
        self.y, self.obs_b = np.asarray(y, dtype=float), obs_b
        self.calls = self.points = 0
        q0 = as_csc(build_prior(np.asarray(theta0, dtype=float)))
        n = q0.shape[0]
        self.n = n
        pats = [q0, sp.identity(n, format="csc")]
        if A is not None:
            A = as_csc(A)
            pats.append(sp.csc_matrix(A.T @ A))
        self.kappa_aug = None
        if constraints is not None:
            ca = np.asarray(constraints[0], dtype=float).reshape(-1, n)
            self.kappa_aug = sp.csc_matrix(HARD_CONSTRAINT_KAPPA * (ca.T @ ca))  # constraints.rs:275-313
            pats.append(self.kappa_aug)
        self.indptr, self.indices = _union_pattern(pats)
        self.kappa_vals = (_on_pattern(self.kappa_aug, self.indptr, self.indices)
                           if self.kappa_aug is not None else None)
        oa = np.ones(len(y)) if slot is not None else obs_a
        self.obs_a = oa
        self.model = LaplaceModel(n, self.indptr, self.indices, _on_pattern(q0, self.indptr, self.indices)[None], y,
                                  family=family, link=link, A=A, obs_a=oa, obs_b=obs_b, constraints=constraints,
                                  capacity=capacity)
        self.capacity = capacity

    def qvals(self, th):
        v = _on_pattern(as_csc(self.build_prior(np.asarray(th, dtype=float))), self.indptr, self.indices)
        return v + self.kappa_vals if self.kappa_vals is not None else v

    def dqvals(self, th):
        out = np.empty((self.m, len(self.indices)))
        for l in range(self.m):
            h = 1e-5 * max(1.0, abs(th[l]))
            tp, tm = np.array(th, dtype=float), np.array(th, dtype=float)
            tp[l] += h
            tm[l] -= h
            out[l] = (self.qvals(tp) - self.qvals(tm)) / (2 * h)
        return out

    def log_prior(self, th):
        v = float(self.log_prior_density(np.asarray(th, dtype=float)))
        g = np.zeros(self.m)
        for l in range(self.m):
            h = 1e-5 * max(1.0, abs(th[l]))
            tp, tm = np.array(th, dtype=float), np.array(th, dtype=float)
            tp[l] += h
            tm[l] -= h
            g[l] = (self.log_prior_density(tp) - self.log_prior_density(tm)) / (2 * h)
        return v, g

    def lpar(self, th):
        if self.slot is None:
            return None, None
        lp = np.array([math.exp(t[self.slot]) if self.slot < len(t) else 1.0 for t in th])
        dl = np.zeros((len(th), self.m))
        if self.slot < self.m:
            dl[:, self.slot] = 1.0
        return lp, dl

    def values(self, thetas, grad=False, want_x=False, want_var=False, want_eta=False):
        thetas = np.atleast_2d(np.asarray(thetas, dtype=float))
        acc = {k: [] for k in ("g", "grad", "x", "var", "eta", "eta_var", "status")}
        for i in range(0, len(thetas), self.capacity):
            th = thetas[i:i + self.capacity]
            qv = np.stack([self.qvals(t) for t in th])
            dq = np.stack([self.dqvals(t) for t in th]) if grad else None
            lp, dl = self.lpar(th)
            r = self.model.eval(qvals=qv, dqvals=dq, lpar=lp, dlog_lpar=dl if grad else None, want_x=want_x,
                                want_var=want_var, want_eta=want_eta)
            self.calls += 1
            self.points += len(th)
            pri = [self.log_prior(t) for t in th] if grad else [(float(self.log_prior_density(t)), None) for t in th]
            ok = r["status"] == 0
            ml = r["mlik"].cpu().numpy()
            # a failed Newton loop: 1e12 for the simplex (hyper_opt.rs:39-45), an empty node
            # for the design (inference.rs:1347-1366)
            acc["g"].append(np.where(ok, ml + np.array([p[0] for p in pri]), -1e12))
            acc["status"].append(r["status"])
            if grad:
                acc["grad"].append(r["grad"].cpu().numpy() + np.array([p[1] for p in pri]))
            for k, fill in (("x", 0.0), ("var", 1.0), ("eta", 0.0), ("eta_var", 1.0)):
                if k in r:
                    v = r[k].cpu().numpy()
                    v[~ok] = fill
                    acc[k].append(v)
        return {k: np.concatenate(v) for k, v in acc.items() if v}

    def posterior_precision(self, theta, x):
        """Q_post at theta and x (posterior_precision_at, inference.rs:1027-1058):
        Q (+ the constraint augmentation) + A^T diag(-hess(A x)) A."""
        q = as_csc(self.build_prior(np.asarray(theta, dtype=float)))
        if self.kappa_aug is not None:
            q = sp.csc_matrix(q + self.kappa_aug)
        A = self.model.A
        eta = A @ x if A is not None else x
        lp, _ = self.lpar(np.asarray(theta, dtype=float)[None])
        _, _, w = MS.obs_eval(self.family, self.link, eta[None], self.y, self.obs_a, self.obs_b, lp)
        w = w.cpu().numpy()[0]
        At = A if A is not None else sp.identity(len(x), format="csc")
        out = sp.csc_matrix(q + At.T @ sp.diags(w) @ At)
        out.sort_indices()
        return out


# --------------------------------------------------------------------------- helpers
def _union_pattern(mats):
    n = mats[0].shape[0]
    pat = sp.csc_matrix((n, n))
    for mtx in mats:
        mtx = as_csc(mtx)
        pat = pat + sp.csc_matrix((np.ones(mtx.nnz), mtx.indices, mtx.indptr), shape=mtx.shape)
    pat = sp.csc_matrix(pat)
    pat.sum_duplicates()
    pat.sort_indices()
    return pat.indptr.astype(np.int64), pat.indices.astype(np.int64)


def _on_pattern(mtx, indptr, indices):
    from .workloads import values_on_pattern

    return values_on_pattern(mtx, indptr, indices)


def gaussian_mixture_marginal(means, vars_, weights, n_points, n_sd=_N_SD):
    """marginals.rs:85-128: the weighted Gaussian mixture on a uniform grid over the
    mixture mean +- n_sd mixture sd, renormalised by the trapezoid rule."""
    means, vars_, weights = (np.asarray(v, dtype=float) for v in (means, vars_, weights))
    if not (len(means) == len(vars_) == len(weights)) or len(means) == 0:
        raise InvalidArgument("mixture means/vars/weights length mismatch")
    if n_points < 5:
        raise InvalidArgument("n_points must be >= 5")
    mu = float(weights @ means)
    var = float(weights @ (vars_ + (means - mu) ** 2))
    if not (var > 0.0 and math.isfinite(var)):
        var = 1e-8
    half = max(n_sd, 1.0) * math.sqrt(var)
    x = mu - half + 2.0 * half * np.arange(n_points) / (n_points - 1)
    good = (vars_ > 0.0) & np.isfinite(vars_)
    s = np.sqrt(np.where(good, vars_, 1.0))
    z = (x[:, None] - means[None]) / s[None]
    dens = np.where(good[None], np.exp(-0.5 * z * z) / (s[None] * math.sqrt(2.0 * math.pi)), 0.0)
    y = dens @ weights
    mass = float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(x)))
    if mass > 0.0 and math.isfinite(mass):
        y = y / mass
    return Marginal1D(x, y)


def hyperpar_marginals(theta_nodes, weights, n_points, n_sd=_N_SD):
    """marginals.rs:130-180: per theta_j a Gaussian kernel mixture at the nodes with the
    weighted sample variance (floored at 1e-8) as bandwidth."""
    out = []
    for j in range(theta_nodes.shape[1]):
        col = theta_nodes[:, j]
        mu = float(weights @ col)
        h2 = max(float(weights @ (col - mu) ** 2), 1e-8)
        out.append(gaussian_mixture_marginal(col, np.full(len(col), h2), weights, n_points, n_sd))
    return out


def marginal_log_lik_gaussian(log_post_at_mode, neg_hessian):
    """compute_marginal_log_lik_gaussian (model_selection.rs:25-61)."""
    m = neg_hessian.shape[0]
    if m == 0:
        return float(log_post_at_mode)
    try:
        L = np.linalg.cholesky(neg_hessian)
    except np.linalg.LinAlgError:
        return float("nan")
    log_det = 2.0 * float(np.sum(np.log(np.diag(L))))
    return float(log_post_at_mode + 0.5 * m * math.log(2.0 * math.pi) - 0.5 * log_det)


# --------------------------------------------------------------------------- driver
def integrate(nodes, initial_theta, strategy="ccd", step_or_f0=1.0, n_points=201, latent_marginal_indices=None,
              predictor_marginal_indices=None, dic=True, waic=True, cpo=True, mode_search="nelder_mead",
              lik=None):
    """The node loop and aggregation of run_inla_inference_a_cancellable
    (inference.rs:1246-1597) over a node evaluator (GpuNodes): mode search, Hessian
    stencil, design, one batched evaluation of the nodes, mixture, model selection and
    marginals. ``lik`` = (family, link, y, obs_a, obs_b) for DIC / WAIC / CPO."""
    theta0 = np.asarray(initial_theta, dtype=float)
    m = len(theta0)
    if mode_search == "bfgs":
        mode, _ = F.bfgs_mode(nodes, theta0)
    elif mode_search == "nelder_mead":
        mode = F.nelder_mead_mode(nodes, theta0)  # hyper_opt.rs:49-66, step 0.1, 200 iterations, tol 1e-6
    else:
        raise InvalidArgument("mode_search must be nelder_mead or bfgs")
    h = F._hessian_step(mode)
    stencil = ccd.hessian_points(mode, h)
    H = ccd.hessian_from_values(nodes.values(stencil)["g"], m, h)
    thetas, z_w = F._nodes(mode, H, strategy.lower(), step_or_f0, lambda th: nodes.values(th)["g"])
    lam, V = ccd.scaling_from_hessian(H)
    z = np.array([np.linalg.solve(V * np.sqrt(np.abs(lam))[None, :], t - mode) for t in thetas])
    mode_index = int(np.argmin(np.linalg.norm(z, axis=1)))
    r = nodes.values(thetas, want_x=True, want_var=True, want_eta=True)
    lp = r["g"].copy()
    lp[r["status"] != 0] = -np.inf
    if not np.any(np.isfinite(lp)):
        raise RuntimeError("Newton-Raphson did not converge at any integration node")
    mx = float(np.max(lp))
    dens = np.exp(lp - mx)
    s = float(np.sum(z_w * dens))
    w = z_w * dens / s
    xm = w @ r["x"]
    xv = w @ (r["var"] + (r["x"] - xm) ** 2)
    em = w @ r["eta"]
    ev = w @ (r["eta_var"] + (r["eta"] - em) ** 2)
    out = dict(mode=mode, hessian=H, latent_means=xm, latent_variances=xv, predictor_means=em,
               predictor_variances=ev, marginal_log_lik=mx + math.log(abs(s)),
               marginal_log_lik_gaussian=marginal_log_lik_gaussian(lp[mode_index], -H),
               theta_nodes=thetas, node_weights=w, log_posts=lp)
    nan = float("nan")
    fam, link, y, oa, ob = lik
    lpm = nodes.lpar(mode[None])[0]
    lp_k = None if lpm is None else np.full(len(thetas), lpm[0])  # the mode's observations
    oa_eff = np.ones(len(y)) if lpm is not None else oa
    if dic:
        d = MS.compute_dic(fam, link, y, r["eta"], w, mode_index, oa_eff, ob, lp_k)
        out.update(dic=d["dic"], mean_deviance=d["mean_deviance"], effective_params=d["effective_params"])
    else:
        out.update(dic=nan, mean_deviance=nan, effective_params=nan)
    if waic:
        wr = MS.compute_waic(fam, link, y, r["eta"], w, oa_eff, ob, lp_k)
        out.update(waic=wr["waic"], waic_lppd=wr["lppd"], waic_effective_params=wr["effective_params"])
    else:
        out.update(waic=nan, waic_lppd=nan, waic_effective_params=nan)
    if cpo:
        c = MS.compute_cpo_pit(fam, link, y, r["eta"], r["eta_var"], w, oa_eff, ob, lp_k)
        out.update(cpo=[None if np.isnan(v) else float(v) for v in c["cpo"]],
                   pit=[None if np.isnan(v) else float(v) for v in c["pit"]], cpo_n_failures=c["n_failures"])
    else:
        out.update(cpo=[None] * len(y), pit=[None] * len(y), cpo_n_failures=0)
    li = list(latent_marginal_indices or [])
    pi = list(predictor_marginal_indices or [])
    for idx in li:
        if not 0 <= idx < r["x"].shape[1]:
            raise InvalidArgument(f"latent marginal index {idx} out of range (n={r['x'].shape[1]})")
    for idx in pi:
        if not 0 <= idx < len(y):
            raise InvalidArgument(f"predictor marginal index {idx} out of range (n_obs={len(y)})")
    out.update(
        internal_marginals_hyperpar=hyperpar_marginals(thetas, w, n_points) if m > 0 else [],
        marginals_latent=[gaussian_mixture_marginal(r["x"][:, i], r["var"][:, i], w, n_points) for i in li],
        marginals_latent_indices=li,
        marginals_predictor=[gaussian_mixture_marginal(r["eta"][:, i], r["eta_var"][:, i], w, n_points) for i in pi],
        marginals_predictor_indices=pi,
        mode_index=mode_index, x_nodes=r["x"])
    return out


def run_inla_inference(initial_theta, build_prior, log_prior_density, obs, strategy="ccd", step_or_f0=1.0,
                       n_points=201, latent_marginal_indices=None, predictor_marginal_indices=None, a=None,
                       constraints_a=None, constraints_e=None, deterministic=False, gaussian_free_prec=False,
                       dic=True, waic=True, cpo=True, mode_search="nelder_mead", capacity=16) -> InferenceResult:
    """py-inla's run_inla_inference (py-inla/src/lib.rs:427-657) on the GPU.

    The arguments are the reference's. ``deterministic`` is accepted: the device
    reductions run in a fixed order, so results do not depend on it. Two additions:
    - ``mode_search``: "nelder_mead" (default, the reference's search, one point per
      call) or "bfgs" (the adjoint gradient, with batched line-search points);
    - ``capacity``: the largest batch of θ-points per evaluation.
    Returns an InferenceResult with the reference's fields."""
    del deterministic
    theta0 = np.asarray(initial_theta, dtype=float)
    fam, link, y, oa, ob, slot = parse_obs(obs, gaussian_free_prec)
    constraints = None
    if (constraints_a is None) != (constraints_e is None):
        raise InvalidArgument("constraints_a and constraints_e must both be provided or both omitted")
    if constraints_e is not None:
        ce = np.asarray(constraints_e, dtype=float)
        if len(ce) == 0:
            raise InvalidArgument("constraints_e must be non-empty")
        ca = np.asarray(constraints_a, dtype=float)
        if ca.size % len(ce):
            raise InvalidArgument("constraints_a length must be divisible by len(constraints_e)")
        constraints = (ca.reshape(len(ce), -1), ce)
    nodes = GpuNodes(build_prior, log_prior_density, len(theta0), fam, link, y, oa, ob, slot, A=a,
                     constraints=constraints, capacity=capacity, theta0=theta0)
    r = integrate(nodes, theta0, strategy, step_or_f0, n_points, latent_marginal_indices, predictor_marginal_indices,
                  dic, waic, cpo, mode_search, lik=(fam, link, y, oa, ob))
    qpost = nodes.posterior_precision(r["mode"], r["x_nodes"][r["mode_index"]])
    keys = [f for f in InferenceResult.__dataclass_fields__ if f not in ("posterior_precision", "evaluations", "extra")]
    res = InferenceResult(**{k: r[k] for k in keys}, posterior_precision=qpost, evaluations=nodes.points)
    res.extra = {"batched_calls": nodes.calls, "mode_index": r["mode_index"]}
    return res
