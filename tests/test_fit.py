"""Mode-search drivers of inla_gpu.fit on a host objective (no GPU): BFGS with
batched line search and the Nelder-Mead restatement find a known optimum."""

import numpy as np
import pytest


class QuadObjective:
    """g(theta) = -1/2 (theta - c)^T A (theta - c) + 7, batched like fit.Objective."""

    def __init__(self, A, c):
        self.A, self.c = np.asarray(A, float), np.asarray(c, float)
        self.calls = 0
        self.points = 0

    def values(self, thetas, grad=False, want_x=False, want_var=False):
        th = np.atleast_2d(thetas)
        d = th - self.c
        self.calls += 1
        self.points += len(th)
        out = {"g": -0.5 * np.einsum("bi,ij,bj->b", d, self.A, d) + 7.0}
        if grad:
            out["grad"] = -d @ self.A
        return out


A = np.array([[4.0, 1.0, 0.2], [1.0, 3.0, 0.5], [0.2, 0.5, 2.0]])
C = np.array([1.5, -0.7, 2.2])


def test_bfgs_finds_quadratic_optimum():
    from inla_gpu import fit as F

    ob = QuadObjective(A, C)
    x, it = F.bfgs_mode(ob, np.zeros(3), gtol=1e-10, xtol=1e-12)
    assert np.max(np.abs(x - C)) < 1e-8
    assert it < 30


def test_bfgs_caps_step_and_batches_line_search():
    from inla_gpu import fit as F

    ob = QuadObjective(A * 1e5, C)  # steep: the first step must be capped
    x, _ = F.bfgs_mode(ob, np.zeros(3), gtol=1e-6, xtol=1e-12, max_step=0.5)
    assert np.max(np.abs(x - C)) < 1e-6
    # every line-search call evaluates n_steps points at once
    assert ob.points >= 4 * (ob.calls - 1)


def test_nelder_mead_restatement_converges():
    from inla_gpu import fit as F

    ob = QuadObjective(A, C)
    x = F.nelder_mead_mode(ob, np.zeros(3), step=0.1, max_iter=2000, tol=1e-12)
    assert np.max(np.abs(x - C)) < 1e-4
    assert ob.points == ob.calls  # one theta-point per evaluation


def test_nelder_mead_iteration_cap():
    from inla_gpu import fit as F

    ob = QuadObjective(A, C)
    F.nelder_mead_mode(ob, np.zeros(3), step=0.1, max_iter=5, tol=0.0)
    # m+1 initial vertices, then at most m+2 evaluations per iteration (shrink)
    assert ob.calls <= 4 + 5 * 4


def test_mixture_of_a_single_node():
    from inla_gpu import ccd

    w, mean, var, mlik = ccd.aggregate(np.array([1.0]), np.array([-3.0]), np.array([[1.0, 2.0]]),
                                       np.array([[0.5, 0.25]]))
    assert np.allclose(w, [1.0]) and np.allclose(mean, [1.0, 2.0]) and np.allclose(var, [0.5, 0.25])
    assert mlik == pytest.approx(-3.0)
