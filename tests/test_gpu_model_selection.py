"""DIC / WAIC over CCD nodes (model_selection.rs:62-199) on device log-likelihoods.

The per-observation log-likelihood is pinned against the oracle's eval_lik for every
family; DIC and WAIC through what the definitions fix: one node has no effective
parameters, weight concentrated on one node reproduces that node, duplicated nodes
change nothing, and the Gaussian deviance has a closed form.
"""

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = [
    ("gaussian", "identity", lambda r, m: r.normal(size=m), lambda r, m: r.uniform(0.5, 2.0, m), None),
    ("poisson", "log", lambda r, m: r.poisson(3.0, m).astype(float), lambda r, m: r.uniform(0.5, 2.0, m), None),
    ("binomial", "logit", lambda r, m: r.binomial(5, 0.4, m).astype(float), lambda r, m: np.full(m, 5.0), None),
    ("nbinomial", "log", lambda r, m: r.poisson(4.0, m).astype(float), lambda r, m: r.uniform(0.5, 2.0, m),
     lambda r, m: r.uniform(1.0, 5.0, m)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_obs_loglik_matches_oracle(case):
    from inla_gpu.model_selection import obs_loglik

    fam, link, gy, ga, gb = case
    rng = np.random.default_rng(3)
    m, k = 257, 3
    y, a = gy(rng, m), ga(rng, m)
    b = gb(rng, m) if gb else None
    eta = rng.normal(scale=0.7, size=(k, m)) + (1.0 if link == "log" else 0.0)
    got = obs_loglik(fam, link, eta, y, a, b).cpu().numpy()
    for q in range(k):
        ref, _, _ = orc.eval_lik(fam, link, eta[q], y, a, b)
        assert np.max(np.abs(got[q] - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_single_node_gaussian_closed_form():
    from inla_gpu.model_selection import compute_dic, compute_waic

    rng = np.random.default_rng(5)
    m = 400
    y = rng.normal(size=m)
    a = rng.uniform(0.5, 2.0, m)
    eta = rng.normal(scale=0.3, size=(1, m))
    tau = 1.7
    # N(y; eta, 1 / (tau a)): D = sum tau a (y - eta)^2 - log(tau a) + log(2 pi)
    dev = float(np.sum(tau * a * (y - eta[0]) ** 2 - np.log(tau * a) + np.log(2 * np.pi)))
    d = compute_dic("gaussian", "identity", y, eta, [1.0], 0, obs_a=a, lpar=[tau])
    assert abs(d["mean_deviance"] - dev) <= 1e-10 * abs(dev)
    assert d["effective_params"] == 0.0 and d["dic"] == d["mean_deviance"]
    w = compute_waic("gaussian", "identity", y, eta, [1.0], obs_a=a, lpar=[tau])
    assert abs(w["effective_params"]) <= 1e-12
    assert abs(w["waic"] - dev) <= 1e-10 * abs(dev)


def test_weights_select_and_duplicate_nodes():
    from inla_gpu.model_selection import compute_dic, compute_waic

    rng = np.random.default_rng(7)
    m = 300
    y = rng.poisson(2.0, m).astype(float)
    a = np.ones(m)
    eta = rng.normal(scale=0.4, size=(3, m)) + 0.7
    one = compute_waic("poisson", "log", y, eta[1:2], [1.0], obs_a=a)
    conc = compute_waic("poisson", "log", y, eta, [0.0, 1.0, 0.0], obs_a=a)
    assert abs(conc["waic"] - one["waic"]) <= 1e-10 * abs(one["waic"])
    assert abs(conc["effective_params"]) <= 1e-12
    dup = compute_waic("poisson", "log", y, np.vstack([eta[1], eta[1]]), [0.3, 0.7], obs_a=a)
    assert abs(dup["waic"] - one["waic"]) <= 1e-10 * abs(one["waic"])
    mixed = compute_waic("poisson", "log", y, eta, [0.2, 0.5, 0.3], obs_a=a)
    assert mixed["effective_params"] > 0.0
    # DIC: D-bar is the weighted node deviance; p_D = D-bar - D(mode node)
    dn = [compute_dic("poisson", "log", y, eta[q:q + 1], [1.0], 0, obs_a=a)["mean_deviance"] for q in range(3)]
    d = compute_dic("poisson", "log", y, eta, [0.2, 0.5, 0.3], 1, obs_a=a)
    dbar = 0.2 * dn[0] + 0.5 * dn[1] + 0.3 * dn[2]
    assert abs(d["mean_deviance"] - dbar) <= 1e-10 * abs(dbar)
    assert abs(d["effective_params"] - (dbar - dn[1])) <= 1e-9 * abs(dbar)


def test_argument_errors():
    # model_selection.rs:107-112 and 157-166
    from inla_gpu.model_selection import compute_dic, compute_waic
    from inla_gpu.solver import InvalidArgument

    y = np.zeros(4)
    eta = np.zeros((2, 4))
    with pytest.raises(InvalidArgument):
        compute_dic("gaussian", "identity", y, eta, [1.0], 0)
    with pytest.raises(InvalidArgument):
        compute_dic("gaussian", "identity", y, eta, [0.5, 0.5], 2)
    with pytest.raises(InvalidArgument):
        compute_waic("gaussian", "identity", y, eta, [1.0])


INVALID = [
    ("binomial", "logit", [3.0], [2.0], None),      # y > n
    ("binomial", "logit", [-1.0], [2.0], None),     # y < 0
    ("poisson", "log", [-1.0], [1.0], None),        # negative count
    ("poisson", "log", [1.0], [-1.0], None),        # negative exposure
    ("gaussian", "identity", [0.5], [0.0], None),   # precision <= 0
    ("nbinomial", "log", [1.0], [1.0], [0.0]),      # size <= 0
    ("nbinomial", "log", [1.0], [0.0], [2.0]),      # mean <= 0 (zero exposure)
]


@pytest.mark.parametrize("case", INVALID, ids=[f"{c[0]}-{i}" for i, c in enumerate(INVALID)])
def test_obs_loglik_invalid_inputs_are_errors(case):
    # eval_likelihood returns Err for these (inference.rs:170-311); the device path
    # reports GMRF_INVALID_ARGUMENT instead of a number
    from inla_gpu.model_selection import obs_loglik
    from inla_gpu.solver import InvalidArgument

    fam, link, y, a, b = case
    eta = np.array([[0.3, 0.3]])
    with pytest.raises(InvalidArgument):
        obs_loglik(fam, link, eta, y * 2, a * 2, None if b is None else b * 2)


def test_obs_loglik_zero_poisson_rate_is_minus_infinity():
    # a zero rate with y > 0 is Ok(-inf) in the reference (inference.rs:204-210), not an error
    from inla_gpu.model_selection import obs_loglik

    got = obs_loglik("poisson", "log", np.array([[0.1, 0.1]]), [2.0, 0.0], [0.0, 0.0]).cpu().numpy()
    assert got[0, 0] == -np.inf and got[0, 1] == 0.0
