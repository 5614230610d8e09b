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


# ------------------------------------------------------------------ integration designs
def test_ccd_design_matches_oracle():
    from inla_gpu import ccd
    from oracle import oracle as orc

    for m in range(1, 6):
        for f0 in (1.0, 2.5):
            z, w = ccd.ccd_design(m, f0)
            zo, wo = orc.ccd_design(m, f0)
            assert np.array_equal(z, zo) and np.array_equal(w, wo)


def test_m5_nodes_match_oracle_eigen_step():
    # CCD nodes of a 5-parameter model: (-H)^-1 by Gauss-Jordan, Jacobi eigenvectors
    # (order and signs fix which half fraction of the design the nodes land on)
    from inla_gpu import ccd
    from oracle import oracle as orc

    rng = np.random.default_rng(3)
    G = rng.standard_normal((5, 5))
    H = -(G @ G.T + 5 * np.eye(5))
    lam, V = ccd.scaling_from_hessian(H)
    sig = orc.invert_symmetric(-H)
    lo, Vo = orc.jacobi_eigen(sig)
    assert np.max(np.abs(lam - lo)) < 1e-14 * np.max(np.abs(lo))
    assert np.max(np.abs(V - Vo)) < 1e-13
    mode = rng.standard_normal(5)
    z, _ = ccd.ccd_design(5, 1.0)
    t = np.array([ccd.theta_from_z(mode, lam, V, zz) for zz in z])
    to = np.array([mode + (Vo * np.sqrt(np.abs(lo))) @ zz for zz in z])
    assert np.max(np.abs(t - to)) < 1e-13
    # and the decomposition is one of Sigma
    assert np.max(np.abs(V @ np.diag(lam) @ V.T - np.linalg.inv(-H))) < 1e-13


@pytest.mark.parametrize("m,step", [(1, 1.0), (1, 0.6), (2, 1.0), (2, 0.7), (3, 1.0)])
def test_grid_design_matches_oracle(m, step):
    # a skewed log-density (not quadratic): the walks stop at different distances on
    # the two sides and the 2-D box holds points beyond the threshold
    from inla_gpu import ccd
    from oracle import oracle as orc

    def g(z):
        z = np.atleast_1d(z)
        return -0.5 * z @ z - 0.05 * z[0] ** 4 + 0.3 * z[0] + (0.3 * z[0] * z[1] if len(z) > 1 else 0.0)

    pts, w = ccd.grid_design(m, step, 4.0, lambda Z: np.array([g(zz) for zz in np.atleast_2d(Z)]))
    if m > 2:
        zo, wo = orc.ccd_design(m, 2.5)
    else:
        zo, wo = orc.grid_design(m, step, 4.0, g)
    assert np.array_equal(pts, zo) and np.array_equal(w, wo)


def test_fit_grid_strategy_on_a_quadratic():
    # the grid design through the fit driver: for g = -1/2 (t - c)^T A (t - c) + 7 the
    # log posterior at z is -|z|^2 / 2 + 7, so the 1-D grid is -2..2 and the
    # normalised weights are proportional to exp(-z^2 / 2)
    from inla_gpu import fit as F

    a = np.array([[2.0]])
    c = np.array([0.4])
    ob = QuadObjective(a, c)
    H = -a
    thetas, w = F._nodes(c, H, "grid", 1.0, lambda th: ob.values(th)["g"])
    z = (thetas[:, 0] - c[0]) * np.sqrt(2.0)
    assert np.allclose(z, [0, 1, 2, -1, -2], atol=1e-12) and np.all(w == 1.0)
