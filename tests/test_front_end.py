"""The run_inla_inference front end's host pieces (no GPU): the observation-dict contract
of parse_obs (py-inla/src/lib.rs:1115-1260) and the marginal grids of
crates/inla_stats/src/marginals.rs checked by what fixes them: a one-component mixture is
the normal density (scipy.stats.norm), every grid integrates to one by the trapezoid
rule, and the Gaussian mlik approximation of a quadratic log posterior is the exact
Gaussian integral."""

import math

import numpy as np
import pytest
import scipy.stats as ss


def test_parse_obs_contract():
    from inla_gpu import inference as I
    from inla_gpu.solver import InvalidArgument

    fam, link, y, a, b, slot = I.parse_obs([{"family": "negative_binomial", "y": 2, "exposure": 1.5, "size": 4}])
    assert (fam, link, slot) == ("nbinomial", "log", None) and a[0] == 1.5 and b[0] == 4
    fam, link, *_ = I.parse_obs([{"family": "binomial", "y": 1, "n": 3, "link": "logit"}])
    assert (fam, link) == ("binomial", "logit")
    with pytest.raises(InvalidArgument):
        I.parse_obs([{"family": "poisson", "y": 1, "E": 1}, {"family": "binomial", "y": 1, "n": 2}])
    with pytest.raises(InvalidArgument):
        I.parse_obs([{"family": "poisson", "y": 1, "E": 1, "link": "probit"}])
    with pytest.raises(InvalidArgument):
        I.parse_obs([None])


def test_marginal_grids():
    from inla_gpu import inference as I

    m = I.gaussian_mixture_marginal([0.7], [0.25], [1.0], 201)
    assert len(m.x) == 201 and abs(m.x[0] - (0.7 - 4 * 0.5)) < 1e-14 and abs(m.x[-1] - (0.7 + 4 * 0.5)) < 1e-14
    assert np.allclose(m.y, ss.norm.pdf(m.x, 0.7, 0.5) / np.trapezoid(ss.norm.pdf(m.x, 0.7, 0.5), m.x), rtol=1e-12)
    m2 = I.gaussian_mixture_marginal([0.0, 2.0], [1.0, 0.5], [0.3, 0.7], 101)
    assert abs(np.trapezoid(m2.y, m2.x) - 1.0) < 1e-12
    mu = 0.3 * 0.0 + 0.7 * 2.0
    var = 0.3 * (1.0 + mu ** 2) + 0.7 * (0.5 + (2.0 - mu) ** 2)
    assert abs(m2.x[50] - mu) < 1e-12 and abs(m2.x[-1] - m2.x[0] - 8 * math.sqrt(var)) < 1e-12
    hp = I.hyperpar_marginals(np.array([[0.0, 1.0], [0.5, 1.0], [-0.5, 1.0]]), np.array([0.5, 0.25, 0.25]), 51)
    assert len(hp) == 2 and all(abs(np.trapezoid(h.y, h.x) - 1.0) < 1e-12 for h in hp)
    with pytest.raises(Exception):
        I.gaussian_mixture_marginal([0.0], [1.0], [1.0], 4)


def test_gaussian_mlik_approximation_is_exact_for_a_quadratic():
    from inla_gpu import inference as I

    rng = np.random.default_rng(1)
    g = rng.standard_normal((3, 3))
    P = g @ g.T + 3 * np.eye(3)      # -H of log p(theta) = c - theta^T P theta / 2
    c = -1.7
    exact = c + 0.5 * 3 * math.log(2 * math.pi) - 0.5 * np.linalg.slogdet(P)[1]  # log of the integral
    assert abs(I.marginal_log_lik_gaussian(c, P) - exact) < 1e-12
    assert math.isnan(I.marginal_log_lik_gaussian(c, -P))
