"""Integration over theta: CCD design, finite-difference Hessian stencil, aggregation.

* ``ccd_design`` restates crates/inla_math/src/integration.rs:1-62.
* ``hessian_points`` / ``hessian_from_values`` restate the stencil and formulas
  of crates/inla_math/src/hyper_opt.rs:197-259, laid out as one batch so every
  stencil point is evaluated in a single batched Laplace call.
* ``theta_from_z`` is the node mapping of crates/inla_stats/src/inference.rs:1318-1325.
* ``aggregate`` is the mixture of inference.rs:1419-1468.
"""

from __future__ import annotations

import math

import numpy as np


def ccd_design(m: int, f0: float = 1.0):
    if m == 0:
        raise ValueError("CCD requires at least 1 factor")
    pts = [np.zeros(m)]
    for i in range(m):
        p1 = np.zeros(m)
        p1[i] = f0
        p2 = np.zeros(m)
        p2[i] = -f0
        pts += [p1, p2]
    if 2 <= m <= 4:
        val = f0 / math.sqrt(m)
        for k in range(1 << m):
            pts.append(np.array([val if (k >> i) & 1 else -val for i in range(m)]))
    elif m == 5:
        val = f0 / math.sqrt(5.0)
        for k in range(16):
            p = np.zeros(5)
            sp = 1.0
            for i in range(4):
                s = 1.0 if (k >> i) & 1 else -1.0
                p[i] = s * val
                sp *= s
            p[4] = sp * val
            pts.append(p)
    elif m > 5:
        raise ValueError(f"CCD strategy for m={m} is not supported")
    n_exp = len(pts)
    w = 1.0 / ((n_exp - 1) * (1.0 + math.exp(-0.5 * f0 * f0) * (f0 * f0 / m - 1.0))) if n_exp > 1 else 1.0
    weights = np.full(n_exp, w)
    weights[0] = 1.0 - (n_exp - 1) * w
    return np.array(pts), weights


def hessian_points(mode, h):
    """Stencil of hyper_opt.rs:214-255: mode, +/-h e_i, and the 4 mixed points per i<j."""
    mode = np.asarray(mode, dtype=float)
    m = len(mode)
    pts = [mode.copy()]
    for i in range(m):
        for s in (h, -h):
            p = mode.copy()
            p[i] += s
            pts.append(p)
    for i in range(m):
        for j in range(i + 1, m):
            for si, sj in ((h, h), (h, -h), (-h, h), (-h, -h)):
                p = mode.copy()
                p[i] += si
                p[j] += sj
                pts.append(p)
    return np.array(pts)


def hessian_from_values(g, m, h):
    """Hessian of g (= log posterior) from values at ``hessian_points``."""
    H = np.zeros((m, m))
    g0 = g[0]
    for i in range(m):
        H[i, i] = (g[1 + 2 * i] - 2 * g0 + g[2 + 2 * i]) / (h * h)
    k = 1 + 2 * m
    for i in range(m):
        for j in range(i + 1, m):
            gpp, gpm, gmp, gmm = g[k:k + 4]
            H[i, j] = H[j, i] = (gpp - gpm - gmp + gmm) / (4 * h * h)
            k += 4
    return H


def scaling_from_hessian(H):
    """Eigen-decomposition of Sigma = (-H)^-1 (inference.rs:1267-1270)."""
    sigma = np.linalg.inv(-H)
    lam, V = np.linalg.eigh(sigma)
    return lam, V


def theta_from_z(mode, lam, V, z):
    return np.asarray(mode) + (V * np.sqrt(np.abs(lam))) @ np.asarray(z)


def aggregate(z_weights, log_post, means, variances):
    """Normalised weights, mixture means and variances (inference.rs:1419-1468)."""
    log_post = np.asarray(log_post)
    mx = np.max(log_post)
    dens = np.exp(log_post - mx)
    s = np.sum(z_weights * dens)
    w = z_weights * dens / s
    mean = w @ means
    var = w @ (variances + (means - mean) ** 2)
    mlik = mx + math.log(abs(s))
    return w, mean, var, mlik
