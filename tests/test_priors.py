"""Hyperparameter priors of the benchmark configurations (host, no GPU), against the
reference's own test values: pc.cor1(0.5, 0.75) at theta = 0 is -2.381562305990987
(py-inla/tests/test_pc_priors.py:70-75, rel 1e-8); pc.prec is the closed form of
priors.rs:532-538 (its test priors.rs:950-960). Gradients against central
differences."""

import math

import pytest


def test_pc_cor1_reference_value():
    from inla_gpu import workloads as W

    lam = W.pc_cor1_lambda(0.5, 0.75)
    assert W.pc_cor1_prior(0.0, lam)[0] == pytest.approx(-2.381562305990987, rel=1e-8)
    # the defining tail probability: P(rho > u) = alpha
    assert (1 - math.exp(-lam * math.sqrt(0.5))) / (1 - math.exp(-lam * math.sqrt(2.0))) == pytest.approx(0.75, abs=1e-13)


def test_pc_prec_closed_form():
    from inla_gpu import workloads as W

    lam = -math.log(0.01)
    for th in (-2.0, 0.0, 1.3):
        v, _ = W.pc_prec_prior(th)
        assert v == pytest.approx(math.log(lam / 2.0) - lam * math.exp(-th / 2.0) - th / 2.0, rel=1e-14)


@pytest.mark.parametrize("th", [-1.5, 0.2, 2.5])
def test_prior_gradients(th):
    from inla_gpu import workloads as W

    lam = W.pc_cor1_lambda(0.5, 0.75)
    h = 1e-5
    for f in (lambda t: W.pc_prec_prior(t), lambda t: W.pc_cor1_prior(t, lam)):
        fd = (f(th + h)[0] - f(th - h)[0]) / (2 * h)
        assert f(th)[1] == pytest.approx(fd, rel=1e-7, abs=1e-9)
