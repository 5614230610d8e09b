"""Linear-combination summaries on the GPU factor of the posterior precision
(crates/inla_stats/src/posterior.rs:53-91, the binding's lincomb of
py-inla/src/lib.rs:339-366): mean a^T mu, sd = sqrt(a^T Q^-1 a), 1.96 sd bounds, for
combinations with repeated indices, against the dense inverse and the oracle's solve;
the reference's errors (index out of range, non-finite weight, Q of the wrong size);
and through InferenceResult.lincomb after a front-end run."""

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _qpost():
    from inla_gpu import workloads as W

    wl = W.spde_gaussian_workload(nx=18, n_obs=900, seed=4)
    c, _, lp, _ = wl.params(wl.theta_true[None])
    q = sp.csc_matrix(wl.gmrf.csc(c[0]) + lp[0] * (wl.A.T @ wl.A))
    q.sum_duplicates()
    q.sort_indices()
    return q


def test_lincomb_matches_dense_inverse_and_oracle(record_error):
    from inla_gpu.posterior import lincomb_summaries

    q = _qpost()
    n = q.shape[0]
    rng = np.random.default_rng(3)
    mu = rng.standard_normal(n)
    combs = [("single", [(5, 1.0)]),
             ("contrast", [(0, 1.0), (n - 1, -1.0)]),
             ("repeated", [(7, 0.5), (7, 0.5), (40, 2.0)]),
             ("dense", [(int(i), float(w)) for i, w in zip(rng.choice(n, 60, replace=False), rng.standard_normal(60))])]
    combs += [(f"c{k}", [(int(i), 1.0) for i in rng.choice(n, 5, replace=False)]) for k in range(300)]
    out = lincomb_summaries(mu, q, combs, chunk=128)
    assert [o["name"] for o in out] == [c[0] for c in combs]
    qinv = np.linalg.inv(q.toarray())
    ref = orc.Factor(q)
    err_sd, err_m = 0.0, 0.0
    for (name, ws), o in zip(combs, out):
        a = np.zeros(n)
        for i, w in ws:
            a[i] += w
        sd = math.sqrt(a @ qinv @ a)
        sd_o = math.sqrt(a @ ref.solve(a))
        err_sd = max(err_sd, abs(o["sd"] - sd) / sd, abs(o["sd"] - sd_o) / sd_o)
        err_m = max(err_m, abs(o["mean"] - a @ mu) / max(abs(a @ mu), 1e-300))
        assert o["q50"] == o["mean"]
        assert o["q025"] == pytest.approx(o["mean"] - 1.96 * o["sd"], rel=1e-15, abs=1e-15)
        assert o["q975"] == pytest.approx(o["mean"] + 1.96 * o["sd"], rel=1e-15, abs=1e-15)
    record_error("lincomb sd vs dense inverse and oracle", err_sd, 1e-10)
    assert err_sd < 1e-10 and err_m < 1e-12


def test_lincomb_errors_are_the_references():
    from inla_gpu.posterior import lincomb_summaries
    from inla_gpu.solver import InvalidArgument

    q = _qpost()
    n = q.shape[0]
    mu = np.zeros(n)
    with pytest.raises(InvalidArgument, match=f"lincomb 'bad': index {n} out of range"):
        lincomb_summaries(mu, q, [("ok", [(0, 1.0)]), ("bad", [(n, 1.0)])])
    with pytest.raises(InvalidArgument, match="lincomb 'nan': non-finite weight"):
        lincomb_summaries(mu, q, [("nan", [(1, float("nan"))])])
    with pytest.raises(InvalidArgument, match="latent mean has length"):
        lincomb_summaries(mu[:-1], q, [("ok", [(0, 1.0)])])
    assert isinstance(InvalidArgument("x"), ValueError)


def test_inference_result_lincomb():
    from inla_gpu import inference as I
    from inla_gpu import workloads as W

    wl = W.spde_gaussian_workload(nx=10, n_obs=300, seed=21)
    g = wl.gmrf

    def build_prior(theta):
        c, _ = W.spde_coefs(theta[1], theta[2])
        return g.csc(c)

    def log_prior(theta):
        return wl.log_prior(theta)[0]

    obs = [{"family": "gaussian", "y": v, "precision": 1.0, "prec_theta": 0} for v in wl.y]
    res = I.run_inla_inference(list(wl.theta_true), build_prior, log_prior, obs, strategy="ccd", a=wl.A,
                               gaussian_free_prec=True, capacity=16)
    out = res.lincomb([("x3", [(3, 1.0)]), ("sum", [(0, 1.0), (1, 1.0)])])
    qinv = np.linalg.inv(res.posterior_precision.toarray())
    assert out[0]["mean"] == pytest.approx(res.latent_means[3], rel=1e-14)
    assert out[0]["sd"] == pytest.approx(math.sqrt(qinv[3, 3]), rel=1e-10)
    assert out[1]["sd"] == pytest.approx(math.sqrt(qinv[0, 0] + qinv[1, 1] + 2 * qinv[0, 1]), rel=1e-10)
