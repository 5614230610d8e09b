"""DIC, WAIC and CPO/PIT over the CCD nodes (crates/inla_stats/src/model_selection.rs:62-511).

The per-observation log-likelihoods log pi(y_i | eta_ki) and their derivatives come
from the device likelihood kernel (gmrf_obs_loglik / gmrf_obs_eval, the same code as the
Laplace evaluation); the node sums follow compute_dic / compute_waic / compute_cpo_pit.
The node predictors and their variances are the Laplace evaluation's eta / eta_var
outputs (LaplaceModel.eval(want_eta=True)).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from .laplace import _dev
from .solver import InvalidArgument, check

FAMILIES = {"gaussian": 0, "poisson": 1, "binomial": 2, "nbinomial": 3}
LINKS = {"identity": 0, "log": 1, "logit": 2}


def obs_loglik(family, link, eta, y, obs_a=None, obs_b=None, lpar=None) -> torch.Tensor:
    """log pi(y_i | eta_ki) for eta of shape (K, n_obs); a (K, n_obs) float64 cuda tensor.

    obs_a: Gaussian precision scale / Poisson exposure / binomial trials / NB exposure
    (None = 1); obs_b: NB size; lpar: (K,) Gaussian precision multipliers (None = 1).
    """
    if family not in FAMILIES or link not in LINKS:
        raise InvalidArgument(f"unknown family or link: {family}, {link}")
    e = _dev(eta)
    if e.dim() != 2:
        raise InvalidArgument("eta must be (nodes, n_obs)")
    k, m = e.shape
    yv = _dev(y, (m,))
    av = _dev(np.ones(m) if obs_a is None else obs_a, (m,))
    bv = _dev(obs_b, (m,)) if obs_b is not None else None
    lp = _dev(lpar, (k,)) if lpar is not None else None
    out = torch.empty((k, m), dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.lib.gmrf_obs_loglik(FAMILIES[family], LINKS[link], k, m, e.data_ptr(), yv.data_ptr(),
                                av.data_ptr(), bv.data_ptr() if bv is not None else None,
                                lp.data_ptr() if lp is not None else None, out.data_ptr(), st))
    return out


def obs_eval(family, link, eta, y, obs_a=None, obs_b=None, lpar=None):
    """(logp, grad, neg_hess), each (K, n_obs) float64 cuda tensors: log pi(y_i | eta_ki),
    its first derivative in eta and minus its second (the Newton weight), as
    eval_likelihood returns them (inference.rs:170-311)."""
    if family not in FAMILIES or link not in LINKS:
        raise InvalidArgument(f"unknown family or link: {family}, {link}")
    e = _dev(eta)
    if e.dim() != 2:
        raise InvalidArgument("eta must be (nodes, n_obs)")
    k, m = e.shape
    yv = _dev(y, (m,))
    av = _dev(np.ones(m) if obs_a is None else obs_a, (m,))
    bv = _dev(obs_b, (m,)) if obs_b is not None else None
    lp = _dev(lpar, (k,)) if lpar is not None else None
    out = [torch.empty((k, m), dtype=torch.float64, device="cuda") for _ in range(3)]
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.lib.gmrf_obs_eval(FAMILIES[family], LINKS[link], k, m, e.data_ptr(), yv.data_ptr(), av.data_ptr(),
                              bv.data_ptr() if bv is not None else None, lp.data_ptr() if lp is not None else None,
                              out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), st))
    return tuple(out)


def compute_dic(family, link, y, cond_predictors, norm_weights, mode_index, obs_a=None, obs_b=None,
                lpar=None) -> dict:
    """model_selection.rs:101-134: D-bar = sum_k w_k D(eta_k), p_D = D-bar - D(eta_mode),
    DIC = D-bar + p_D, with D(eta) = -2 sum_i log pi(y_i | eta_i)."""
    w = np.asarray(norm_weights, dtype=np.float64)
    eta = _dev(cond_predictors)
    if eta.dim() != 2 or len(w) != eta.shape[0]:
        raise InvalidArgument("norm_weights length must match cond_predictors length")
    if not 0 <= mode_index < eta.shape[0]:
        raise InvalidArgument("mode_index out of range")
    ll = obs_loglik(family, link, eta, y, obs_a, obs_b, lpar).cpu().numpy()
    dev = -2.0 * ll.sum(axis=1)
    mean_dev = float(np.dot(w, dev))
    p_d = mean_dev - float(dev[mode_index])
    return {"dic": mean_dev + p_d, "mean_deviance": mean_dev, "effective_params": p_d}


def compute_waic(family, link, y, cond_predictors, norm_weights, obs_a=None, obs_b=None,
                 lpar=None) -> dict:
    """model_selection.rs:152-199: lppd_i = log sum_k w_k exp(l_ik) (max-shifted),
    p_i = sum_k w_k (l_ik - sum_k w_k l_ik)^2, WAIC = -2 (sum lppd_i - sum p_i)."""
    w = np.asarray(norm_weights, dtype=np.float64)
    eta = _dev(cond_predictors)
    if eta.dim() != 2 or len(w) != eta.shape[0]:
        raise InvalidArgument("norm_weights length must match cond_predictors length")
    ll = obs_loglik(family, link, eta, y, obs_a, obs_b, lpar).cpu().numpy()  # (K, n)
    mx = ll.max(axis=0)
    s = np.einsum("k,ki->i", w, np.exp(ll - mx))
    if not np.all(np.isfinite(s) & (s > 0.0)):
        i = int(np.flatnonzero(~(np.isfinite(s) & (s > 0.0)))[0])
        raise InvalidArgument(f"WAIC: non-finite predictive density at obs {i}")
    lppd = float(np.sum(mx + np.log(s)))
    mean = np.einsum("k,ki->i", w, ll)
    p = float(np.sum(np.einsum("k,ki->i", w, (ll - mean) ** 2)))
    return {"waic": -2.0 * (lppd - p), "lppd": lppd, "effective_params": p}


CPO_DIFF = 3.0                          # model_selection.rs (CPO_DIFF)
_CPO_STEPS = np.array([-3.0, -1.5, 0.0, 1.5, 3.0])


def normal_cdf(x):
    """The normal CDF formula of model_selection.rs:450-465: the Abramowitz-Stegun 7.1.26
    rational form, applied with exp(-x^2 / 2) as the reference writes it (it is
    1/2 (1 + sign(x) (1 - erfc(|x|) exp(x^2 / 2))) up to 1.5e-7 exp(x^2 / 2), not Phi)."""
    a1, a2, a3, a4, a5, p = 0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.3275911
    x = np.asarray(x, dtype=np.float64)
    ax = np.abs(x)
    t = 1.0 / (1.0 + p * ax)
    y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * np.exp(-ax * ax / 2.0)
    return 0.5 * (1.0 + np.where(x < 0.0, -1.0, 1.0) * y)


def _poisson_cdf(lam, k):
    """Pr(X <= k), X ~ Poisson(lam), elementwise by the direct summation of
    model_selection.rs:470-485 (the same terms in the same order)."""
    lam = np.asarray(lam, dtype=np.float64)
    k = np.broadcast_to(np.asarray(k, dtype=np.int64), lam.shape)
    cdf = np.zeros(lam.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        lp = -lam
        loglam = np.log(lam)
        for j in range(int(k.max(initial=0)) + 1):
            on = j <= k
            cdf = np.where(on, cdf + np.exp(lp), cdf)
            lp = np.where(j < k, lp + loglam - math.log(j + 1), lp)
    return np.where(lam <= 0.0, 1.0, np.minimum(cdf, 1.0))


def _binomial_cdf(n, p, k):
    """Pr(X <= k), X ~ Binomial(n, p), elementwise by the direct summation of
    model_selection.rs:488-503."""
    n = np.asarray(n, dtype=np.int64)
    p = np.asarray(p, dtype=np.float64)
    shape = np.broadcast_shapes(n.shape, p.shape, np.shape(k))
    n, p, k = np.broadcast_to(n, shape), np.broadcast_to(p, shape), np.broadcast_to(np.asarray(k, np.int64), shape)
    from scipy.special import gammaln

    cdf = np.zeros(shape)
    lgn = gammaln(n + 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        lp, lq = np.log(p), np.log(1.0 - p)
        for j in range(int(k.max(initial=0)) + 1):
            on = j <= k
            lc = lgn - math.lgamma(j + 1.0) - gammaln(np.maximum(n - j, 0) + 1.0)
            cdf = np.where(on, cdf + np.exp(lc + j * lp + (n - j) * lq), cdf)
    out = np.minimum(cdf, 1.0)
    out = np.where(p >= 1.0, np.where(k >= n, 1.0, 0.0), out)
    return np.where(p <= 0.0, 1.0, out)


def _observation_cdf(family, eta, variance, y, a):
    """observation_cdf (model_selection.rs:398-447): NaN where the reference has None
    (negative binomial). The link is identity / log / logit as the reference assumes."""
    if family == "gaussian":
        tv = 1.0 / a + variance
        with np.errstate(invalid="ignore", divide="ignore"):
            out = normal_cdf((y - eta) / np.sqrt(tv))
        return np.where(tv <= 0.0, np.nan, out)
    if family == "poisson":
        lam = a * np.exp(eta)
        yi = np.floor(y).astype(np.int64)
        ok = (lam >= 0.0) & np.isfinite(lam)
        out = np.where(yi < 0, 0.0, _poisson_cdf(np.where(ok, lam, 1.0), np.maximum(yi, 0)))
        return np.where(ok, out, np.nan)
    if family == "binomial":
        pr = np.where(eta >= 0.0, 1.0 / (1.0 + np.exp(-np.abs(eta))), np.exp(-np.abs(eta)) / (1.0 + np.exp(-np.abs(eta))))
        n = np.floor(a).astype(np.int64)
        yi = np.floor(y).astype(np.int64)
        out = _binomial_cdf(n, pr, np.clip(yi, 0, n))
        return np.where(yi < 0, 0.0, out)
    return np.full(np.shape(eta), np.nan)


def compute_cpo_pit(family, link, y, cond_predictors, cond_pred_vars, norm_weights, obs_a=None, obs_b=None,
                    lpar=None) -> dict:
    """compute_cpo_pit (model_selection.rs:222-325): for each observation the failure
    heuristic of check_cpo_failure_for_obs at every node (leave-one-out precision
    1/var - c, the unnormalised leave-one-out log density at five points, monotonicity
    and the CPO_DIFF border test), then CPO_i = 1 / sum_k w_k / pi(y_i | eta_ki) and
    PIT_i = sum_k w_k F(y_i | eta_ki, var_ki). The likelihood, its gradient and c = -hess
    at the node predictors, and at the 5 K leave-one-out points, are evaluated on the
    device in two calls. lpar: (K,) Gaussian precision multipliers of the observations
    (the reference evaluates every node with the mode's observations: pass the mode's
    multiplier K times). Returns cpo and pit (NaN where the reference returns None) and
    n_failures."""
    w = np.asarray(norm_weights, dtype=np.float64)
    eta = np.asarray(cond_predictors.cpu().numpy() if isinstance(cond_predictors, torch.Tensor) else cond_predictors,
                     dtype=np.float64)
    var = np.asarray(cond_pred_vars.cpu().numpy() if isinstance(cond_pred_vars, torch.Tensor) else cond_pred_vars,
                     dtype=np.float64)
    if eta.ndim != 2 or var.shape != eta.shape or len(w) != eta.shape[0]:
        raise InvalidArgument("cond_predictors, cond_pred_vars, and norm_weights must have equal length")
    K, n = eta.shape
    y = np.asarray(y, dtype=np.float64)
    a = np.ones(n) if obs_a is None else np.broadcast_to(np.asarray(obs_a, dtype=np.float64), (n,))
    lp_k = np.ones(K) if lpar is None else np.asarray(lpar, dtype=np.float64).reshape(K)
    logp, grad, c = (t.cpu().numpy() for t in obs_eval(family, link, eta, y, a, obs_b, lp_k))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        fail = ~((var > 0.0) & np.isfinite(var))
        tau_loo = 1.0 / var - c
        fail |= ~((tau_loo > 0.0) & np.isfinite(tau_loo))
        var_loo = 1.0 / tau_loo
        sd = np.sqrt(var_loo)
        mu_loo = var_loo * (eta / var - c * eta + grad)
        X = mu_loo[None] + _CPO_STEPS[:, None, None] * sd[None]            # (5, K, n)
        X = np.where(fail[None] | ~np.isfinite(X), eta[None], X)             # failed: any finite point
    ll5 = obs_loglik(family, link, X.reshape(5 * K, n), y, a, obs_b, np.tile(lp_k, 5)).cpu().numpy()
    with np.errstate(invalid="ignore"):
        d = -0.5 * (X - eta[None]) ** 2 / var[None] + ll5.reshape(5, K, n)
        inc = np.all(d[1:] >= d[:-1], axis=0)
        dec = np.all(d[1:] <= d[:-1], axis=0)
        mx = d.max(axis=0)
        fail |= inc | dec | (mx - d[0] < CPO_DIFF) | (mx - d[4] < CPO_DIFF) | ~np.isfinite(mx)
    failed = fail.any(axis=0) | (K < 2)
    with np.errstate(over="ignore", invalid="ignore"):
        max_neg = (-logp).max(axis=0)
        inv = (w[:, None] * np.exp(-logp - max_neg[None])).sum(axis=0) * np.exp(max_neg)
    bad_inv = ~(np.isfinite(inv) & (inv > 0.0))
    cpo_ok = ~failed & ~bad_inv
    cpo = np.where(cpo_ok, 1.0 / np.where(cpo_ok, inv, 1.0), np.nan)
    f = _observation_cdf(family, eta, var, y[None, :], a[None, :] * (lp_k[:, None] if family == "gaussian" else 1.0))
    ps = (w[:, None] * f).sum(axis=0)
    pit = np.where(cpo_ok & np.isfinite(ps), np.clip(ps, 0.0, 1.0), np.nan)
    return {"cpo": cpo, "pit": pit, "n_failures": int(np.sum(~cpo_ok))}
