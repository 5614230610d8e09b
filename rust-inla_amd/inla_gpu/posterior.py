"""Post-inference linear combinations on the GPU factor of the posterior precision.

lincomb_summaries follows crates/inla_stats/src/posterior.rs:53-91: for each
combination v = a^T x, mean a^T mu from the mixture mean and Var(v) = a^T Q^-1 a at
the posterior precision stored at the mode (the Gaussian-at-mode approximation), with
the 1.96 sd bounds. The reference factorises Q once (FaerCpuSolver, posterior.rs:68-69)
and solves one combination at a time; here Q is factorised once on the device and all
combination vectors are solved in multi-right-hand-side launch sequences
(gmrf_batch_solve_multi) of up to ``chunk`` vectors (fewer for large n: at most 256 MB
of dense vectors per sequence).
"""

from __future__ import annotations

import math

import numpy as np

from .solver import Batch, InvalidArgument, Symbolic, as_csc, check


def _dense_vector(name, weights, n):
    """LinComb::dense_vector (posterior.rs:35-50): duplicate indices are summed."""
    a = np.zeros(n)
    for i, w in weights:
        i = int(i)
        if i < 0 or i >= n:
            raise InvalidArgument(f"lincomb '{name}': index {i} out of range [0, {n})")
        w = float(w)
        if not math.isfinite(w):
            raise InvalidArgument(f"lincomb '{name}': non-finite weight")
        a[i] += w
    return a


def lincomb_summaries(mean, q_post, combs, chunk=256):
    """Gaussian summaries of linear combinations (posterior.rs:53-91).

    mean: latent mixture mean (length n); q_post: the posterior precision (n x n CSC,
    both triangles); combs: ``[(name, [(latent_index, weight), ...]), ...]``. Returns a
    list of dicts with name, mean, sd, q025, q50, q975, as the Python binding's
    ``lincomb`` (py-inla/src/lib.rs:339-366). Errors are ``InvalidArgument``
    (a ``ValueError``), as the binding's ``PyValueError``; a failed factorisation raises
    the solver's error.
    """
    import torch

    mean = np.ascontiguousarray(mean, dtype=np.float64)
    n = len(mean)
    q = as_csc(q_post)
    if q.shape != (n, n):
        raise InvalidArgument(f"lincomb: Q is {q.shape[0]}×{q.shape[1]}, latent mean has length {n}")
    sym = Symbolic(n, q.indptr, q.indices)
    bt = Batch(sym, 1)
    bt.factorize(torch.as_tensor(q.data[None, :], dtype=torch.float64, device="cuda"))
    _, status = bt.logdet()
    check(int(status[0]))
    # at most 2^25 doubles (256 MB) of dense combination vectors per solve
    chunk = max(1, min(int(chunk), (1 << 25) // max(n, 1)))
    out = []
    for c0 in range(0, len(combs), chunk):
        part = combs[c0:c0 + chunk]
        A = np.stack([_dense_vector(name, weights, n) for name, weights in part])
        w = bt.solve_multi(torch.as_tensor(A[None], dtype=torch.float64, device="cuda"))[0].cpu().numpy()
        for (name, _), a, wi in zip(part, A, w):
            m = float(a @ mean)
            var = float(a @ wi)
            sd = math.sqrt(max(var, 0.0))
            out.append({"name": name, "mean": m, "sd": sd, "q025": m - 1.96 * sd, "q50": m,
                        "q975": m + 1.96 * sd})
    return out
