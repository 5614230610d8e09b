"""DIC and WAIC over the CCD nodes (crates/inla_stats/src/model_selection.rs:62-199).

The per-observation log-likelihoods log pi(y_i | eta_ki) come from the device
likelihood kernel (gmrf_obs_loglik, the same code as the Laplace evaluation); the
node sums follow compute_dic / compute_waic.
"""

from __future__ import annotations

import ctypes as C

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
