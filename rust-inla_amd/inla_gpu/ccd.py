"""Integration over theta: CCD and grid designs, finite-difference Hessian stencil,
eigen-scaling, aggregation.

* ``ccd_design`` restates crates/inla_math/src/integration.rs:1-62.
* ``grid_design`` restates integration.rs:64-156 with the evaluator called on
  batches: each step asks for the points it needs at once, so one batched Laplace
  call covers a whole line of the grid.
* ``invert_symmetric`` and ``jacobi_eigen`` restate integration.rs:158-226 and
  249-318 (the reference's own Gauss-Jordan and Jacobi sweeps: deterministic
  eigenvector signs and order, which the m = 5 half fraction depends on).
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


def invert_symmetric(h):
    """Gauss-Jordan with partial pivoting on [A | I] (integration.rs:158-226)."""
    h = np.asarray(h, dtype=float)
    m = h.shape[0]
    a = np.hstack([h, np.eye(m)])
    for i in range(m):
        # the first row of largest magnitude, as the reference's strict > scan
        best = abs(a[i, i])
        mr = i
        for r in range(i + 1, m):
            if abs(a[r, i]) > best:
                best, mr = abs(a[r, i]), r
        if abs(a[mr, i]) < 1e-14:
            raise np.linalg.LinAlgError("Singular Hessian matrix")
        if mr != i:
            a[[i, mr]] = a[[mr, i]]
        a[i] /= a[i, i]
        for r in range(m):
            if r != i and a[r, i] != 0.0:
                a[r] -= a[r, i] * a[i]
    return a[:, m:].copy()


def jacobi_eigen(a, max_iter=100):
    """Cyclic Jacobi rotations (integration.rs:249-318): eigenvalues and the
    eigenvectors as columns of V, in the order and signs the sweeps produce."""
    a = np.array(a, dtype=float)
    m = a.shape[0]
    v = np.eye(m)
    for _ in range(max_iter):
        off = sum(abs(a[i, j]) for i in range(m) for j in range(i + 1, m))
        if off < 1e-12:
            return np.diag(a).copy(), v
        for p in range(m):
            for q in range(p + 1, m):
                apq = a[p, q]
                if abs(apq) < 1e-15:
                    continue
                app, aqq = a[p, p], a[q, q]
                tau = (aqq - app) / (2.0 * apq)
                t = 1.0 / (tau + math.sqrt(1.0 + tau * tau)) if tau >= 0.0 else \
                    -1.0 / (-tau + math.sqrt(1.0 + tau * tau))
                c = 1.0 / math.sqrt(1.0 + t * t)
                sn = t * c
                a[p, p] = app - t * apq
                a[q, q] = aqq + t * apq
                a[p, q] = a[q, p] = 0.0
                for r in range(m):
                    if r != p and r != q:
                        arp, arq = a[r, p], a[r, q]
                        a[r, p] = a[p, r] = c * arp - sn * arq
                        a[r, q] = a[q, r] = sn * arp + c * arq
                vp, vq = v[:, p].copy(), v[:, q].copy()
                v[:, p] = c * vp - sn * vq
                v[:, q] = sn * vp + c * vq
    raise np.linalg.LinAlgError("Jacobi eigenvalue algorithm did not converge")


def scaling_from_hessian(H):
    """Eigen-decomposition of Sigma = (-H)^-1 (inference.rs:1267-1270), by the
    reference's inversion and Jacobi sweeps."""
    sigma = invert_symmetric(-np.asarray(H, dtype=float))
    return jacobi_eigen(sigma, 100)


def grid_design(m: int, step: float, threshold: float, evaluate, max_steps: int = 10_000):
    """Grid design of integration.rs:64-156 (m = 1, 2; m > 2 falls back to the CCD
    design with f0 = 2.5 as there). ``evaluate(Z)`` returns g (the log posterior) at
    each row of Z. Points are accepted in the reference's order; the outward walks
    ask for ``chunk`` points per call and keep those before the first one beyond the
    threshold, which gives the same points as the reference's one-at-a-time loop.
    Returns (points, weights = 1)."""
    if m == 0:
        raise ValueError("Grid requires at least 1 factor")
    if m > 2:
        return ccd_design(m, 2.5)
    chunk = 8
    g0 = float(evaluate(np.zeros((1, m)))[0])

    def walk(direction):
        # number of consecutive steps k = 1, 2, ... along `direction` within the threshold
        # (the reference walks without a bound; a log posterior that never falls by the
        # threshold would not end there either: here it is an error after max_steps)
        k = 0
        while k < max_steps:
            ks = np.arange(k + 1, k + 1 + chunk)
            g = np.asarray(evaluate(ks[:, None] * step * np.asarray(direction, dtype=float)[None, :]))
            beyond = np.nonzero(g0 - g > threshold)[0]
            if len(beyond):
                return k + int(beyond[0])
            k += chunk
        raise RuntimeError(f"grid_design: the log posterior does not fall by {threshold} "
                           f"within {max_steps} steps")

    if m == 1:
        kp, km = walk([1.0]), walk([-1.0])
        pts = [[0.0]] + [[k * step] for k in range(1, kp + 1)] + [[-(k * step)] for k in range(1, km + 1)]
    else:
        max_i, min_i = walk([1.0, 0.0]), walk([-1.0, 0.0])
        max_j, min_j = walk([0.0, 1.0]), walk([0.0, -1.0])
        cand = np.array([[i * step, j * step] for i in range(-min_i, max_i + 1)
                         for j in range(-min_j, max_j + 1)])
        g = np.asarray(evaluate(cand))
        pts = cand[g0 - g <= threshold].tolist()
    pts = np.array(pts, dtype=float)
    return pts, np.ones(len(pts))


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
