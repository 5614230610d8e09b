"""Small test matrices built the way the reference's own tests build them."""

import numpy as np
import scipy.sparse as sp


def ar1_q(n, rho, tau):
    """crates/inla_math/src/sparse_ldlt.rs:590-605 (tests) / crates/inla_stats/src/ar1.rs:9-63."""
    main = np.full(n, tau * (1 + rho * rho))
    main[0] = main[-1] = tau
    off = np.full(n - 1, -tau * rho)
    return sp.csc_matrix(sp.diags([off, main, off], [-1, 0, 1]))


def besag_ridge_q(adj, tau, ridge):
    """crates/inla_math/src/sparse_ldlt.rs:607-622 (tests)."""
    n = len(adj)
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i)
        cols.append(i)
        vals.append(tau * len(adj[i]) + ridge)
        for j in adj[i]:
            if j > i:
                rows += [i, j]
                cols += [j, i]
                vals += [-tau, -tau]
    q = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    q.sum_duplicates()
    q.sort_indices()
    return q


FOUR_CYCLE_CHORD = [[1, 3], [0, 2, 3], [1, 3], [0, 1, 2]]
PATH6 = [[j for j in (i - 1, i + 1) if 0 <= j < 6] for i in range(6)]


def mixture_major(n_time, n_comp, diag, clique, chain):
    """FGN-like mixture-major layout (sparse_ldlt.rs:526-547, ordering.rs:169-189)."""
    n = n_time * n_comp
    rows, cols, vals = [], [], []
    for t in range(n_time):
        for c in range(n_comp):
            i = c * n_time + t
            rows.append(i); cols.append(i); vals.append(diag)
            for d in range(c + 1, n_comp):
                j = d * n_time + t
                rows += [i, j]; cols += [j, i]; vals += [clique, clique]
            if t + 1 < n_time and c > 0:
                i2 = c * n_time + t + 1
                rows += [i, i2]; cols += [i2, i]; vals += [chain, chain]
    q = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    q.sum_duplicates()
    q.sort_indices()
    return q


def random_spd(n, density, seed, shift=None):
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=density, random_state=rng, format="csc")
    a = a + a.T
    d = np.asarray(abs(a).sum(axis=0)).ravel() + (1.0 if shift is None else shift)
    q = sp.csc_matrix(a + sp.diags(d))
    q.sum_duplicates()
    q.sort_indices()
    return q


def shuffled(q, seed):
    rng = np.random.default_rng(seed)
    p = rng.permutation(q.shape[0])
    q = sp.csc_matrix(q)[p][:, p]
    q = sp.csc_matrix(q)
    q.sort_indices()
    return q
