"""The low-rank form of the constraint augmentation (host algebra, no GPU).

The reference factorises Q + kappa A_c^T A_c explicitly (constraints.rs:275-313).
The GPU model with constraint_kappa instead grounds Q at one node per constraint
row, Q_g = Q + sum_t alpha_t e_p e_p^T (same pattern), and writes
Q + kappa A_c^T A_c = Q_g + U E U^T, U = [A_c^T, e_p], E = diag(kappa I, -alpha), so
  log|Q_aug| = log|Q_g| + log|E| + log|E^-1 + U^T Q_g^-1 U|,
  Q_aug^-1  = Q_g^-1 - Y K^-1 Y^T,  Y = Q_g^-1 U, K = E^-1 + U^T Y.
These identities are checked here in dense numpy against independent values: the
pseudo-determinant of an intrinsic Besag x AR1 prior whose null space the
constraints span (log|Q + kappa C^T C| = log pdet Q + k log kappa + log|C C^T|),
and the dense inverse at a moderate kappa where the explicit form is well
conditioned. A planted sign error in E fails them.
"""

import math

import numpy as np
import scipy.sparse as sp


def _besag_ar1(nr=6, T=4, tau=2.0, rho=0.7):
    from inla_gpu import workloads as W

    R = W.besag_structure(W.rook_adjacency(nr, nr)).toarray()
    I, M, O = (m.toarray() for m in W.ar1_pieces(T))
    Q = tau * (np.kron(I, R) + rho * rho * np.kron(M, R) - rho * np.kron(O, R))
    ns = nr * nr
    C = np.zeros((T, nr * nr * T))
    for t in range(T):
        C[t, t * ns:(t + 1) * ns] = 1.0 / math.sqrt(ns)
    return Q, C


def _lowrank(Q, C, kappa, e_sign=1.0):
    n = Q.shape[0]
    k = C.shape[0]
    used = np.zeros(n, bool)
    p = []
    for t in range(k):
        order = np.argsort(-np.abs(C[t]), kind="stable")
        j = next(int(j) for j in order if not used[j])
        used[j] = True
        p.append(j)
    alpha = np.array([Q[j, j] if Q[j, j] > 0 else 1.0 for j in p])
    Qg = Q.copy()
    Qg[p, p] += alpha
    U = np.hstack([C.T, np.eye(n)[:, p]])
    E = np.r_[np.full(k, kappa), -e_sign * alpha]
    Y = np.linalg.solve(Qg, U)
    K = np.diag(1.0 / E) + U.T @ Y
    sk, lk = np.linalg.slogdet(K)
    sg, lg = np.linalg.slogdet(Qg)
    logdet = lg + np.sum(np.log(np.abs(E))) + lk
    sign = sg * sk * np.prod(np.sign(E))
    Kinv = np.linalg.inv(K)
    inv = np.linalg.inv(Qg) - Y @ Kinv @ Y.T
    return logdet, sign, inv


def test_prior_logdet_equals_pseudo_determinant():
    Q, C = _besag_ar1()
    kappa = 3.059023205018258e6
    ld, sign, _ = _lowrank(Q, C, kappa)
    ev = np.linalg.eigvalsh(Q)
    k = C.shape[0]
    assert np.all(np.abs(ev[:k]) < 1e-10) and ev[k] > 1e-3  # null space of dimension k
    exact = np.sum(np.log(ev[k:])) + k * math.log(kappa) + np.linalg.slogdet(C @ C.T)[1]
    assert sign > 0
    assert abs(ld - exact) < 1e-10 * abs(exact)


def test_inverse_and_logdet_match_dense_at_moderate_kappa():
    Q, C = _besag_ar1()
    rng = np.random.default_rng(5)
    for kappa, post in ((10.0, False), (10.0, True), (1e3, True)):
        Qb = Q + (np.diag(rng.uniform(0.5, 2.0, Q.shape[0])) if post else 0.0)
        Qa = Qb + kappa * C.T @ C
        ld, sign, inv = _lowrank(Qb, C, kappa)
        assert sign > 0
        assert abs(ld - np.linalg.slogdet(Qa)[1]) < 1e-11 * abs(ld)
        ref = np.linalg.inv(Qa)
        assert np.max(np.abs(inv - ref)) < 1e-10 * np.max(np.abs(ref))
        # trace of the inverse against a derivative direction
        D = sp.random(Q.shape[0], Q.shape[0], density=0.05, random_state=1).toarray()
        D = D + D.T
        assert abs(np.trace(inv @ D) - np.trace(ref @ D)) < 1e-9 * max(1.0, abs(np.trace(ref @ D)))


def test_planted_sign_error_is_caught():
    Q, C = _besag_ar1()
    kappa = 3.059023205018258e6
    ld, sign, _ = _lowrank(Q, C, kappa, e_sign=-1.0)
    ev = np.linalg.eigvalsh(Q)
    k = C.shape[0]
    exact = np.sum(np.log(ev[k:])) + k * math.log(kappa) + np.linalg.slogdet(C @ C.T)[1]
    assert not (sign > 0 and abs(ld - exact) < 1e-6 * abs(exact))
