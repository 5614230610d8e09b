"""Fitting many independent models at once (fit_blocks): every model's BFGS mode search,
Hessian stencil and CCD nodes as block evaluations of one Laplace model, against each
model fitted alone with the same driver (fit_batched on its own LaplaceModel) and
against the oracle's log posterior at the nodes.

The per-model searches see the same values up to rounding (the union changes only the
launch sequence), so they take the same steps: modes agree to 1e-8, Hessians to 1e-5
relative (finite differences of values agreeing to 1e-11), and node weights, mixture
means and variances to 1e-7.
"""

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_fit_blocks_matches_models_fitted_alone(record_error):
    from inla_gpu import fit as F
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel
    from test_gpu_hetero_laplace import small_models

    wl = W.bym2_poisson_batch(models=small_models([70, 160, 45, 230], seed=9))
    k = wl.k
    union = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=9)
    obj = F.BlockObjective(union, wl.params, W.bym2_log_prior)
    theta0 = np.tile([0.5, 0.5], (k, 1))
    fits, secs = F.fit_blocks(obj, theta0, wl.block_ptr)
    assert len(fits) == k and secs["bfgs_rounds"] >= 1
    bp = wl.block_ptr
    for q in range(k):
        e0, e1 = wl.indptr[bp[q]], wl.indptr[bp[q + 1]]
        single = LaplaceModel(bp[q + 1] - bp[q], wl.indptr[bp[q]:bp[q + 1] + 1] - e0, wl.indices[e0:e1] - bp[q],
                              wl.components[:, e0:e1], wl.block_y(q), family="poisson", A=wl.block_A(q),
                              capacity=9)

        def params_q(th, q=q):
            c, dc, lp, dl = wl.params(np.repeat(np.atleast_2d(th)[:, None, :], k, axis=1))
            return c[:, q], dc[:, q], None, None

        alone = F.fit_batched(F.Objective(single, params_q, W.bym2_log_prior), theta0[q])
        fb = fits[q]
        assert np.max(np.abs(fb.mode - alone.mode)) < 1e-8, (q, fb.mode, alone.mode)
        assert np.max(np.abs(fb.hessian - alone.hessian)) < 1e-5 * np.max(np.abs(alone.hessian))
        assert np.max(np.abs(fb.thetas - alone.thetas)) < 1e-6
        e_w = record_error("block fit weights vs alone", float(np.max(np.abs(fb.weights - alone.weights))), 1e-7, "abs")
        e_m = record_error("block fit mean vs alone",
                           float(np.max(np.abs(fb.mean - alone.mean)) / np.max(np.abs(alone.mean))), 1e-7)
        e_v = record_error("block fit var vs alone",
                           float(np.max(np.abs(fb.var - alone.var)) / np.max(np.abs(alone.var))), 1e-7)
        assert e_w < 1e-7 and e_m < 1e-7 and e_v < 1e-7
        # the node log posteriors against the oracle's find_mode on the model alone
        for t, lp in zip(fb.thetas[:3], fb.log_post[:3]):
            o = orc.find_mode(wl.block_prior(q, t), "poisson", "log", wl.block_y(q), 1.0, A=wl.block_A(q))
            e = record_error("block fit node log posterior vs oracle", abs(lp - (o.mlik + W.bym2_log_prior(t)[0])),
                             1e-8, "abs")
            assert e < 1e-8
        # the mode is a stationary point of the log posterior
        g = F.Objective(single, params_q, W.bym2_log_prior).values(fb.mode[None], grad=True)["grad"][0]
        assert np.max(np.abs(g)) < 1e-3
