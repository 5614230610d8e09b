"""Synthetic GMRF workloads, built the way rust-inla builds them.

Everything here produces host-side inputs (CSC patterns, affine precision
components, projectors, observations); nothing here factorises anything.

* ``lattice_mesh`` restates py-inla/python/inla/spde.py:23-63 (regular lattice
  over a rectangle, x fastest, two triangles per cell).
* ``fem_blocks`` restates the mass / stiffness assembly of
  crates/inla_fmesher/src/fmesher.rs:544-629 with identity diffusion
  (c0 lumped area/3 on the diagonal, g1 = e_i . e_j / (4 |T|)).
* ``spde_components`` returns the fixed matrices of
  Q(tau, kappa) = tau^2 (kappa^4 C0 + 2 kappa^2 G1 + G1 C0^-1 G1)
  (crates/inla_stats/src/spde.rs:70-142) on one union pattern, so Q(theta) is
  an affine combination sum_k c_k(theta) Q_k.
* ``lattice_projector`` restates the barycentric projector of
  fmesher.rs:175-219 for points inside the lattice mesh.
* ``besag_components`` restates crates/inla_stats/src/besag.rs:127-152.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

GEOM_EPSILON = 1e-15  # fmesher.rs:5


def lattice_mesh(nx: int, ny: int, xlim=(0.0, 1.0), ylim=(0.0, 1.0)):
    """Vertices (nx*ny, 2) and triangles (2(nx-1)(ny-1), 3); spde.py:23-63."""
    if nx < 2 or ny < 2:
        raise ValueError("nx and ny must be >= 2")
    xs = np.linspace(xlim[0], xlim[1], nx)
    ys = np.linspace(ylim[0], ylim[1], ny)
    xx, yy = np.meshgrid(xs, ys, indexing="xy")
    vertices = np.column_stack([xx.ravel(order="C"), yy.ravel(order="C")])
    j, i = np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    v00 = j * nx + i
    v10 = j * nx + i + 1
    v01 = (j + 1) * nx + i
    v11 = (j + 1) * nx + i + 1
    t1 = np.stack([v00, v10, v01], axis=1)
    t2 = np.stack([v10, v11, v01], axis=1)
    tris = np.empty((2 * t1.shape[0], 3), dtype=np.int64)
    tris[0::2] = t1
    tris[1::2] = t2
    return vertices, tris


def fem_blocks(vertices: np.ndarray, tris: np.ndarray):
    """(c0 diagonal, g1 CSC) with identity diffusion; fmesher.rs:544-629."""
    n = vertices.shape[0]
    s0 = vertices[tris[:, 0]]
    s1 = vertices[tris[:, 1]]
    s2 = vertices[tris[:, 2]]
    e = [s2 - s1, s0 - s2, s1 - s0]  # edge_vec(a, b) = a - b
    area = 0.5 * np.abs((s1[:, 0] - s0[:, 0]) * (s2[:, 1] - s0[:, 1])
                        - (s1[:, 1] - s0[:, 1]) * (s2[:, 0] - s0[:, 0]))
    fa = 0.5 * np.abs(e[0][:, 0] * e[1][:, 1] - e[0][:, 1] * e[1][:, 0])
    fa = np.where(fa <= GEOM_EPSILON, area, fa)
    c0 = np.zeros(n)
    np.add.at(c0, tris.ravel(), np.repeat(area / 3.0, 3))
    rows, cols, vals = [], [], []
    for a in range(3):
        for b in range(3):
            # rotate_perp preserves dot products, so H = I gives e_a . e_b
            eab = e[a][:, 0] * e[b][:, 0] + e[a][:, 1] * e[b][:, 1]
            rows.append(tris[:, a])
            cols.append(tris[:, b])
            vals.append(eab / (4.0 * fa))
    g1 = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(n, n))
    g1.sum_duplicates()
    g1.sort_indices()
    return c0, g1


def union_pattern(mats):
    """Common sorted CSC pattern of several square matrices (+ diagonal)."""
    n = mats[0].shape[0]
    acc = sp.identity(n, format="csc") * 0.0
    pat = sp.csc_matrix((np.ones(n), (np.arange(n), np.arange(n))), shape=(n, n))
    for m in mats:
        m = sp.csc_matrix(m)
        pat = pat + sp.csc_matrix((np.ones(m.nnz), m.indices, m.indptr), shape=m.shape)
    pat = sp.csc_matrix(pat)
    pat.sum_duplicates()
    pat.sort_indices()
    del acc
    return pat.indptr.astype(np.int64), pat.indices.astype(np.int64)


def values_on_pattern(m, indptr, indices):
    """Values of sparse matrix m at the given CSC pattern (0 where m has none)."""
    m = sp.csc_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    n = len(indptr) - 1
    indptr = np.asarray(indptr, dtype=np.int64)
    # entries by column-major key col * n + row: sorted in a sorted CSC pattern
    pkey = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr)) * n + np.asarray(indices, dtype=np.int64)
    mkey = np.repeat(np.arange(n, dtype=np.int64), np.diff(m.indptr)) * n + m.indices.astype(np.int64)
    pos = np.searchsorted(pkey, mkey)
    if np.any(pos >= len(pkey)) or np.any(pkey[np.minimum(pos, len(pkey) - 1)] != mkey):
        raise ValueError("matrix has entries outside the pattern")
    out = np.zeros(len(indices))
    out[pos] = m.data
    return out


@dataclass
class AffineGmrf:
    """Q(theta) = sum_k coef_k(theta) * Q_k on one full symmetric CSC pattern."""

    n: int
    indptr: np.ndarray
    indices: np.ndarray
    components: np.ndarray  # (K, nnz)
    names: list = field(default_factory=list)

    def values(self, coefs):
        return np.asarray(coefs) @ self.components

    def csc(self, coefs):
        return sp.csc_matrix((self.values(coefs), self.indices, self.indptr), shape=(self.n, self.n))


def spde_components(nx: int, ny: int) -> tuple[AffineGmrf, np.ndarray, np.ndarray]:
    """SPDE alpha=2 pieces on the nx x ny lattice: components (C0, G1, G1 C0^-1 G1).

    Q(tau, kappa) = tau^2 kappa^4 C0 + 2 tau^2 kappa^2 G1 + tau^2 G1 C0^-1 G1.
    """
    verts, tris = lattice_mesh(nx, ny)
    c0, g1 = fem_blocks(verts, tris)
    n = verts.shape[0]
    C0 = sp.diags(c0, format="csc")
    GCG = sp.csc_matrix(g1 @ sp.diags(1.0 / c0) @ g1)
    indptr, indices = union_pattern([C0, g1, GCG])
    comps = np.stack([values_on_pattern(C0, indptr, indices),
                      values_on_pattern(g1, indptr, indices),
                      values_on_pattern(GCG, indptr, indices)])
    return AffineGmrf(n, indptr, indices, comps, ["C0", "G1", "G1C0invG1"]), verts, tris


def spde_coefs(log_tau: float, log_kappa: float):
    """Coefficients of (C0, G1, GCG) and their derivatives w.r.t. (log tau, log kappa)."""
    t2 = math.exp(2 * log_tau)
    k2 = math.exp(2 * log_kappa)
    c = np.array([t2 * k2 * k2, 2 * t2 * k2, t2])
    # d/dlog_tau: 2c ; d/dlog_kappa: (4, 2, 0) * c
    dc = np.array([2 * c, [4 * c[0], 2 * c[1], 0.0]])
    return c, dc


def lattice_projector(nx: int, ny: int, pts: np.ndarray, xlim=(0.0, 1.0), ylim=(0.0, 1.0)):
    """Barycentric projector A (n_obs x nx*ny) for points inside the lattice mesh."""
    hx = (xlim[1] - xlim[0]) / (nx - 1)
    hy = (ylim[1] - ylim[0]) / (ny - 1)
    fx = (pts[:, 0] - xlim[0]) / hx
    fy = (pts[:, 1] - ylim[0]) / hy
    i = np.clip(np.floor(fx).astype(np.int64), 0, nx - 2)
    j = np.clip(np.floor(fy).astype(np.int64), 0, ny - 2)
    u = fx - i
    v = fy - j
    v00 = j * nx + i
    v10 = v00 + 1
    v01 = v00 + nx
    v11 = v01 + 1
    lower = (u + v) <= 1.0
    verts = np.where(lower[:, None], np.stack([v00, v10, v01], 1), np.stack([v10, v11, v01], 1))
    w = np.where(lower[:, None], np.stack([1 - u - v, u, v], 1),
                 np.stack([1 - v, u + v - 1, 1 - u], 1))
    m = pts.shape[0]
    rows = np.repeat(np.arange(m), 3)
    keep = np.abs(w.ravel()) > GEOM_EPSILON
    A = sp.csc_matrix((w.ravel()[keep], (rows[keep], verts.ravel()[keep])), shape=(m, nx * ny))
    A.sum_duplicates()
    A.sort_indices()
    return A


def rook_adjacency(nr: int, nc: int):
    """Rook neighbours on an nr x nc lattice, node = r*nc + c."""
    adj = [[] for _ in range(nr * nc)]
    for r in range(nr):
        for c in range(nc):
            i = r * nc + c
            if r > 0:
                adj[i].append(i - nc)
            if c > 0:
                adj[i].append(i - 1)
            if c + 1 < nc:
                adj[i].append(i + 1)
            if r + 1 < nr:
                adj[i].append(i + nc)
    return adj


def besag_structure(adj):
    """Unit-tau Besag structure matrix R (besag.rs:127-152 with tau = 1)."""
    n = len(adj)
    rows, cols, vals = [], [], []
    for i, nb in enumerate(adj):
        rows.append(i)
        cols.append(i)
        vals.append(float(len(nb)) if nb else 1.0)
        for j in nb:
            rows.append(i)
            cols.append(j)
            vals.append(-1.0)
    R = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    R.sum_duplicates()
    R.sort_indices()
    return R


def sample_gmrf(Q: sp.csc_matrix, rng: np.random.Generator) -> np.ndarray:
    """One draw x ~ N(0, Q^-1) through SuperLU in symmetric mode (data generation)."""
    from scipy.sparse.linalg import splu

    n = Q.shape[0]
    lu = splu(sp.csc_matrix(Q), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
              options={"SymmetricMode": True})
    # Pr Q Pc = L U with U = D L^T when no pivoting happened (Pr = Pc^T).
    d = lu.U.diagonal()
    if np.any(d <= 0):
        raise ValueError("sample_gmrf: Q is not positive definite")
    z = rng.standard_normal(n) / np.sqrt(d)
    # Solve U' ... : x_p = L^-T z with L unit lower (U = D L^T  =>  L^T = D^-1 U)
    from scipy.sparse.linalg import spsolve_triangular

    Lt = sp.diags(1.0 / d) @ lu.U
    y = spsolve_triangular(sp.csr_matrix(Lt), z, lower=False, unit_diagonal=True)
    # P Q P^T = L D L^T with (P Q P^T)[perm[i], perm[j]] = Q[i, j]: x = P^T y
    return y[lu.perm_c]


# --------------------------------------------------------------------------
# Priors used by the benchmark configurations (crates/inla_stats/src/priors.rs)
# --------------------------------------------------------------------------
def log_gamma_fn(z):
    return math.lgamma(z)


def loggamma_prior(theta, shape=1.0, rate=5e-5):
    """log pi(theta) and d/dtheta for Gamma(shape, rate) on exp(theta) (priors.rs:836-855)."""
    e = math.exp(theta)
    return shape * math.log(rate) - math.lgamma(shape) + shape * theta - rate * e, shape - rate * e


def pc_prec_prior(theta, u=1.0, alpha=0.01):
    """PC prior of a precision on theta = log tau (priors.rs:532-546) and its gradient:
    log(lambda / 2) - lambda e^(-theta/2) - theta/2 with lambda = -log(alpha) / u."""
    lam = -math.log(alpha) / u
    e = math.exp(-0.5 * theta)
    return math.log(lam) - math.log(2.0) - lam * e - 0.5 * theta, 0.5 * lam * e - 0.5


def pc_cor1_lambda(u=0.5, alpha=0.75):
    """lambda of the PC prior for a correlation with base model rho = 1: the root of
    (1 - e^(-lambda sqrt(1 - u))) / (1 - e^(-lambda sqrt 2)) = alpha (the equation of
    priors.rs:611-660, solved here by bracketing instead of a golden-section search)."""
    from scipy.optimize import brentq

    def f(lam):
        return (1.0 - math.exp(-lam * math.sqrt(1.0 - u))) / (1.0 - math.exp(-lam * math.sqrt(2.0))) - alpha

    return brentq(f, 1e-9, 1e3, xtol=1e-15, rtol=4 * np.finfo(float).eps, maxiter=500)


def pc_cor1_prior(theta, lam):
    """PC prior of rho = 2 / (1 + e^-theta) - 1 with base model rho = 1, on theta
    (priors.rs:662-705), and its gradient."""
    e = math.exp(theta)
    rho = 2.0 * e / (1.0 + e) - 1.0
    s = math.sqrt(1.0 - rho)
    logp = (math.log(lam) - lam * s - math.log(1.0 - math.exp(-math.sqrt(2.0) * lam)) - math.log(2.0 * s)
            + math.log(2.0) + theta - 2.0 * math.log1p(e))
    return logp, (lam * s + 1.0) * (1.0 + rho) / 4.0 - rho


def pc_spde_prior(log_tau, log_kappa, range0=0.1, alpha_r=0.5, sigma0=1.0, alpha_s=0.01, d=2.0):
    """PC prior of the 2-D SPDE in (log tau, log kappa) (priors.rs:809-834) and its gradient."""
    lam1 = -math.log(alpha_r) * range0 ** (d / 2)
    lam2 = -math.log(alpha_s) / sigma0
    log_rho = 0.5 * math.log(8.0) - log_kappa
    log_sigma = -0.5 * math.log(4 * math.pi) - log_tau - log_kappa
    dh = d / 2
    e1 = math.exp(-dh * log_rho)
    s_rho = math.log(lam1 * dh) - dh * log_rho - lam1 * e1
    ds_rho = -dh + lam1 * dh * e1              # d/dlog_rho
    es = math.exp(log_sigma)
    s_sig = math.log(lam2) + log_sigma - lam2 * es
    ds_sig = 1.0 - lam2 * es                   # d/dlog_sigma
    # dlog_rho/dlog_tau = 0, dlog_rho/dlog_kappa = -1; dlog_sigma/d(tau, kappa) = (-1, -1)
    return s_rho + s_sig, np.array([-ds_sig, -ds_rho - ds_sig])


@dataclass
class SpdeGaussianWorkload:
    """SPDE alpha=2 on an nx x nx unit-square lattice, Gaussian observations.

    Internal theta = (log prec_eps, log tau, log kappa).
    """

    gmrf: AffineGmrf
    A: sp.csc_matrix
    y: np.ndarray
    theta_true: np.ndarray
    nx: int

    @property
    def n(self):
        return self.gmrf.n

    def params(self, thetas):
        """(coefs (B,3), dcoefs (B,3,3), lpar (B,), dlog_lpar (B,3)) for theta rows."""
        thetas = np.atleast_2d(thetas)
        B = len(thetas)
        coefs = np.zeros((B, 3))
        dco = np.zeros((B, 3, 3))
        lpar = np.exp(thetas[:, 0])
        dl = np.zeros((B, 3))
        dl[:, 0] = 1.0
        for b, th in enumerate(thetas):
            c, dc = spde_coefs(th[1], th[2])
            coefs[b] = c
            dco[b, 1:] = dc
        return coefs, dco, lpar, dl

    def log_prior(self, theta):
        lp0, g0 = loggamma_prior(theta[0])
        lp1, g12 = pc_spde_prior(theta[1], theta[2])
        return lp0 + lp1, np.r_[g0, g12]


def spde_gaussian_workload(nx=142, n_obs=50_000, seed=20241203, range_=0.3, sigma=1.0, sigma_eps=0.1):
    """The n ~ 20k SPDE Gaussian configuration: true range 0.3, sigma 1, sigma_eps 0.1."""
    g, _, _ = spde_components(nx, nx)
    rng = np.random.default_rng(seed)
    kappa = math.sqrt(8.0) / range_
    tau = 1.0 / (math.sqrt(4 * math.pi) * kappa * sigma)
    c, _ = spde_coefs(math.log(tau), math.log(kappa))
    x = sample_gmrf(g.csc(c), rng)
    A = lattice_projector(nx, nx, rng.uniform(size=(n_obs, 2)))
    y = A @ x + sigma_eps * rng.standard_normal(n_obs)
    theta = np.array([math.log(1.0 / sigma_eps ** 2), math.log(tau), math.log(kappa)])
    return SpdeGaussianWorkload(g, A, y, theta, nx)


# --------------------------------------------------------------------------
# BYM2 + Poisson on random planar graphs (heterogeneous batch)
# --------------------------------------------------------------------------
def delaunay_adjacency(npts: int, rng: np.random.Generator):
    """Neighbour lists of the Delaunay triangulation of npts uniform points in the unit square."""
    from scipy.spatial import Delaunay

    pts = rng.uniform(size=(npts, 2))
    tri = Delaunay(pts)
    nb = [set() for _ in range(npts)]
    for a, b, c in tri.simplices:
        nb[a].update((b, c))
        nb[b].update((a, c))
        nb[c].update((a, b))
    return [sorted(s) for s in nb]


def rw2_structure(n: int) -> sp.csc_matrix:
    """Second-order random-walk structure D2^T D2 (unit precision, equal spacing)."""
    D = sp.diags([np.ones(n - 2), -2.0 * np.ones(n - 2), np.ones(n - 2)], [0, 1, 2], shape=(n - 2, n))
    R = sp.csc_matrix(D.T @ D)
    R.sort_indices()
    return R


def grounded(R: sp.csc_matrix, k: int = 1) -> sp.csc_matrix:
    """R without its last k rows and columns (positive definite when the null space of R
    has dimension k and no null vector vanishes on the last k nodes)."""
    n = R.shape[0]
    G = sp.csc_matrix(R[: n - k, : n - k])
    G.sort_indices()
    return G


def ginv_diag_from_grounded(gdiag, gN, N):
    """diag(R^+) from the grounded inverse G (zero-padded g-inverse of R):
    R^+ = P G P with P = I - N N^T, N an orthonormal basis of the null space, so
    diag(R^+)_i = G_ii - 2 sum_a N_ia (G N)_ia + sum_ab N_ia (N^T G N)_ab N_ib."""
    NtGN = N.T @ gN
    return gdiag - 2.0 * np.sum(N * gN, axis=1) + np.einsum("ia,ab,ib->i", N, NtGN, N)


def null_basis(n: int, kind: str = "constant"):
    """Orthonormal null-space basis: constants (Besag on a connected graph) or
    constants + linear trend (RW2)."""
    cols = [np.ones(n)] if kind == "constant" else [np.ones(n), np.arange(n, dtype=float)]
    Nq, _ = np.linalg.qr(np.column_stack(cols))
    return Nq


def scale_factors(structs, kinds=None):
    """scale_model factors exp(mean log diag(R^+)) (crates/inla_math/src/design.rs:151-195)
    of intrinsic structure matrices, without the dense eigen-decomposition.

    R^+ = P G P where G is the inverse of the grounded matrix (the last k nodes
    removed, zero-padded: a generalised inverse of R) and P projects out the
    k-dimensional null space. G's diagonal comes from one selected inversion and
    G N from k solves, all structures in one heterogeneous batch on the GPU."""
    import torch

    from .solver import MultiBatch

    kinds = kinds or ["constant"] * len(structs)
    Ns = [null_basis(R.shape[0], kd) for R, kd in zip(structs, kinds)]
    gs = [grounded(R, N.shape[1]) for R, N in zip(structs, Ns)]
    kmax = max(N.shape[1] for N in Ns)
    mb = MultiBatch(gs, capacity=1)
    vals = torch.as_tensor(np.concatenate([g.data for g in gs]), dtype=torch.float64, device="cuda")
    mb.factorize(vals[None])
    ld, st = mb.logdet()
    if np.any(st != 0):
        raise ValueError("scale_factors: a grounded structure matrix is singular")
    rhs = np.zeros((kmax, mb.n))
    for q, N in enumerate(Ns):
        a, b = mb.n_off[q], mb.n_off[q + 1]
        rhs[: N.shape[1], a:b] = N[: b - a].T
    sol = np.stack([mb.solve(torch.as_tensor(rhs[a][None], device="cuda"))[0].cpu().numpy()
                    for a in range(kmax)])
    mb.selinv()
    d_all = mb.selinv_diag()[0].cpu().numpy()
    out = []
    for q, (R, N) in enumerate(zip(structs, Ns)):
        n, k = R.shape[0], N.shape[1]
        a, b = mb.n_off[q], mb.n_off[q + 1]
        gdiag = np.r_[d_all[a:b], np.zeros(k)]
        gN = np.vstack([sol[:k, a:b].T, np.zeros((k, k))])
        out.append(math.exp(np.mean(np.log(ginv_diag_from_grounded(gdiag, gN, N)))))
    return np.array(out)


@dataclass
class Bym2Model:
    """One BYM2 (combined-field, besag.rs:192-219) + Poisson areal model: n regions,
    Q(theta) = tau (1 - phi) I + tau phi Q*, Q* the scaled Besag structure."""

    adj: list
    R: sp.csc_matrix
    scale: float
    y: np.ndarray
    indptr: np.ndarray = None
    indices: np.ndarray = None
    components: np.ndarray = None

    def __post_init__(self):
        n = self.R.shape[0]
        I = sp.identity(n, format="csc")
        self.indptr, self.indices = union_pattern([I, self.R])
        self.components = np.vstack([values_on_pattern(I, self.indptr, self.indices),
                                     values_on_pattern(self.R * self.scale, self.indptr, self.indices)])

    @property
    def n(self):
        return self.R.shape[0]

    def coefs(self, tau, phi):
        return np.array([tau * (1.0 - phi), tau * phi])

    def values(self, tau, phi):
        return self.coefs(tau, phi) @ self.components


def bym2_hetero_batch(k=64, nmin=500, nmax=20000, seed=20241204, tau=1.0, phi=0.5, intercept=1.0,
                      sample=True):
    """k independent BYM2 + Poisson models on Delaunay-neighbour graphs with sizes
    uniform in [nmin, nmax]; u drawn from the model's Q, y ~ Poisson(exp(intercept + u))
    (with sample=False the counts stay zero: precision matrices only)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=k)
    adjs = [delaunay_adjacency(int(s), rng) for s in sizes]
    Rs = [besag_structure(a) for a in adjs]
    scales = scale_factors(Rs)
    models = []
    for a, R, c in zip(adjs, Rs, scales):
        m = Bym2Model(a, R, float(c), np.zeros(R.shape[0]))
        if sample:
            Q = sp.csc_matrix((m.values(tau, phi), m.indices, m.indptr), shape=(m.n, m.n))
            u = sample_gmrf(Q, rng)
            m.y = rng.poisson(np.exp(intercept + u)).astype(float)
        models.append(m)
    return models


PHI_CLAMP = 1e-6  # phi = clamp(sigmoid(theta), 1e-6, 1 - 1e-6) (structured.rs:335-337)


@dataclass
class Bym2PoissonBatch:
    """k independent BYM2 + Poisson models as one block-diagonal Laplace model
    (LaplaceModel(blocks=block_ptr)). Block q has the latent (beta_q, u_q), the
    intercept first with the fixed prior precision int_prec, and one observation per
    region, eta_i = beta_q + u_qi, y_i ~ Poisson(exp(eta_i)). theta of each model =
    (log tau, logit phi) and Q_q(theta) = tau (1 - phi) I + tau phi Q*_q on u_q
    (besag.rs:192-219): components (I on u, Q* on u, e_beta e_beta^T)."""

    models: list
    n: int
    indptr: np.ndarray
    indices: np.ndarray
    components: np.ndarray  # (3, nnz)
    A: sp.csc_matrix
    y: np.ndarray
    block_ptr: np.ndarray   # k + 1 latent offsets
    obs_ptr: np.ndarray     # k + 1 observation offsets
    int_prec: float

    @property
    def k(self):
        return len(self.models)

    def params(self, thetas):
        """thetas (B, k, 2) -> (coefs (B, k, 3), dcoefs (B, k, 2, 3), None, None)."""
        th = np.asarray(thetas, dtype=float).reshape(-1, self.k, 2)
        tau = np.exp(th[..., 0])
        s = 1.0 / (1.0 + np.exp(-th[..., 1]))
        phi = np.clip(s, PHI_CLAMP, 1.0 - PHI_CLAMP)
        dphi = np.where((s > PHI_CLAMP) & (s < 1.0 - PHI_CLAMP), s * (1.0 - s), 0.0)
        coefs = np.stack([tau * (1.0 - phi), tau * phi, np.full_like(tau, self.int_prec)], axis=-1)
        dco = np.zeros(th.shape[:2] + (2, 3))
        dco[..., 0, 0], dco[..., 0, 1] = tau * (1.0 - phi), tau * phi
        dco[..., 1, 0], dco[..., 1, 1] = -tau * dphi, tau * dphi
        return coefs, dco, None, None

    def block_prior(self, q, theta) -> sp.csc_matrix:
        """The block's own Q(theta) ((n_q + 1) x (n_q + 1), sorted CSC)."""
        c, _, _, _ = self.params(np.broadcast_to(np.asarray(theta, dtype=float), (self.k, 2))[None])
        a, b = self.block_ptr[q], self.block_ptr[q + 1]
        e0, e1 = self.indptr[a], self.indptr[b]
        ip = self.indptr[a:b + 1] - e0
        v = c[0, q] @ self.components[:, e0:e1]
        Q = sp.csc_matrix((v, self.indices[e0:e1] - a, ip), shape=(b - a, b - a))
        Q.sort_indices()
        return Q

    def block_A(self, q) -> sp.csc_matrix:
        a, b = self.block_ptr[q], self.block_ptr[q + 1]
        o0, o1 = self.obs_ptr[q], self.obs_ptr[q + 1]
        return sp.csc_matrix(self.A[o0:o1, a:b])

    def block_y(self, q) -> np.ndarray:
        return self.y[self.obs_ptr[q]:self.obs_ptr[q + 1]]


def bym2_log_prior(theta):
    """log pi(theta) and its gradient for theta = (log tau, logit phi) of the BYM2 batch:
    loggamma(1, 5e-5) on tau (the C1 prior) and N(0, 1) on logit phi (this workload's
    choice; the reference's BYM2 defaults to a PC prior on phi, which needs the spectrum
    of each graph)."""
    v, g = loggamma_prior(theta[0])
    return v - 0.5 * theta[1] ** 2 - 0.5 * math.log(2 * math.pi), np.array([float(np.ravel(g)[0]), -theta[1]])


def bym2_poisson_batch(k=64, nmin=500, nmax=20000, seed=20241204, tau=1.0, phi=0.5, intercept=1.0,
                       int_prec=1e-3, models=None):
    """The heterogeneous Laplace batch: bym2_hetero_batch's k models (sizes uniform in
    [nmin, nmax], y ~ Poisson(exp(intercept + u))) with an intercept each."""
    models = models if models is not None else bym2_hetero_batch(k, nmin, nmax, seed, tau, phi, intercept)
    ips, ixs, comps, As = [], [], [], []
    bp, op = [0], [0]
    for mdl in models:
        nq = mdl.n
        N = nq + 1
        E0 = sp.csc_matrix(([1.0], ([0], [0])), shape=(N, N))
        Iu = sp.block_diag([sp.csc_matrix((1, 1)), sp.identity(nq)], format="csc")
        Ru = sp.block_diag([sp.csc_matrix((1, 1)), mdl.R * mdl.scale], format="csc")
        Aq = sp.csc_matrix(sp.hstack([sp.csc_matrix(np.ones((nq, 1))), sp.identity(nq)]))
        ip, ix = union_pattern([E0, Iu, Ru, sp.csc_matrix(Aq.T @ Aq)])
        ips.append(ip)
        ixs.append(ix)
        comps.append(np.stack([values_on_pattern(Iu, ip, ix), values_on_pattern(Ru, ip, ix),
                               values_on_pattern(E0, ip, ix)]))
        As.append(Aq)
        bp.append(bp[-1] + N)
        op.append(op[-1] + nq)
    e_off = np.cumsum([0] + [len(ix) for ix in ixs])
    indptr = np.concatenate([np.array([0], dtype=np.int64)] +
                            [ip[1:] + e_off[q] for q, ip in enumerate(ips)]).astype(np.int64)
    indices = np.concatenate([ix + bp[q] for q, ix in enumerate(ixs)]).astype(np.int64)
    A = sp.block_diag(As, format="csc")
    A.sort_indices()
    y = np.concatenate([m.y for m in models])
    return Bym2PoissonBatch(models, bp[-1], indptr, indices, np.concatenate(comps, axis=1), A, y,
                            np.array(bp, dtype=np.int64), np.array(op, dtype=np.int64), int_prec)


@dataclass
class SpdePoissonWorkload:
    """SPDE alpha=2 on an nx x nx lattice plus an intercept, Poisson counts:
    latent x = (beta0, u), eta = beta0 + A u, y ~ Poisson(exp(eta)).
    Internal theta = (log tau, log kappa); beta0 has a fixed prior precision."""

    gmrf: AffineGmrf
    A: sp.csc_matrix
    y: np.ndarray
    theta_true: np.ndarray
    nx: int
    int_prec: float

    @property
    def n(self):
        return self.gmrf.n

    def params(self, thetas):
        """(coefs (B,4), dcoefs (B,2,4), lpar None, dlog_lpar None)."""
        thetas = np.atleast_2d(thetas)
        B = len(thetas)
        coefs = np.zeros((B, 4))
        dco = np.zeros((B, 2, 4))
        for b, th in enumerate(thetas):
            c, dc = spde_coefs(th[0], th[1])
            coefs[b] = np.r_[self.int_prec, c]
            dco[b, :, 1:] = dc
        return coefs, dco, None, None

    def log_prior(self, theta):
        return pc_spde_prior(theta[0], theta[1])


def spde_poisson_workload(nx=142, n_obs=50_000, seed=20241205, range_=0.3, sigma=1.0, intercept=1.0,
                          int_prec=1e-3):
    """The batched Poisson configuration on the SPDE mesh: y ~ Poisson(exp(1 + u))."""
    g, _, _ = spde_components(nx, nx)
    rng = np.random.default_rng(seed)
    kappa = math.sqrt(8.0) / range_
    tau = 1.0 / (math.sqrt(4 * math.pi) * kappa * sigma)
    c, _ = spde_coefs(math.log(tau), math.log(kappa))
    u = sample_gmrf(g.csc(c), rng)
    Au = lattice_projector(nx, nx, rng.uniform(size=(n_obs, 2)))
    y = rng.poisson(np.exp(intercept + Au @ u)).astype(float)
    n = g.n + 1
    A = sp.csc_matrix(sp.hstack([sp.csc_matrix(np.ones((n_obs, 1))), Au]))
    A.sort_indices()
    E0 = sp.csc_matrix(([1.0], ([0], [0])), shape=(n, n))
    comps_q = [sp.block_diag([sp.csc_matrix((1, 1)), g.csc(e)], format="csc") for e in np.eye(3)]
    indptr, indices = union_pattern([E0] + comps_q + [sp.csc_matrix(A.T @ A)])
    comps = np.stack([values_on_pattern(E0, indptr, indices)] +
                     [values_on_pattern(q, indptr, indices) for q in comps_q])
    gm = AffineGmrf(n, indptr, indices, comps, ["intercept", "C0", "G1", "G1C0invG1"])
    return SpdePoissonWorkload(gm, A, y, np.array([math.log(tau), math.log(kappa)]), nx, int_prec)


# --------------------------------------------------------------------------
# Space x time configurations (Kronecker precisions, group-major layout)
# --------------------------------------------------------------------------
def ar1_pieces(T: int):
    """(I, M_inner, O) with the reference AR1 precision (crates/inla_stats/src/ar1.rs:9-63)
    tau * (I + rho^2 M_inner - rho O): diagonal 1 at the ends and 1 + rho^2 inside,
    off-diagonal -rho."""
    I = sp.identity(T, format="csc")
    inner = np.ones(T)
    inner[0] = inner[-1] = 0.0
    M = sp.diags(inner, format="csc")
    O = sp.csc_matrix(sp.diags([np.ones(T - 1), np.ones(T - 1)], [-1, 1]))
    return I, M, O


def rho_of(theta_rho):
    """rho = 2 / (1 + exp(-theta)) - 1, clamped to +-0.999 (structured.rs:353), and drho/dtheta."""
    rho = 2.0 / (1.0 + math.exp(-theta_rho)) - 1.0
    rc = min(max(rho, -0.999), 0.999)
    return rc, (0.5 * (1.0 - rho * rho) if rc == rho else 0.0)


@dataclass
class KronWorkload:
    """A Kronecker-structured latent model with an affine precision, observations,
    a likelihood family and optional linear constraints."""

    gmrf: AffineGmrf
    A: object
    y: np.ndarray
    obs_a: np.ndarray
    obs_b: object
    family: str
    constraints: object
    theta_true: np.ndarray
    params_fn: object
    log_prior_fn: object = None

    @property
    def n(self):
        return self.gmrf.n

    def params(self, thetas):
        return self.params_fn(np.atleast_2d(thetas))

    def log_prior(self, theta):
        return self.log_prior_fn(np.asarray(theta, dtype=float))


def besag_ar1_workload(nr=100, T=20, seed=20241206, tau=2.0, rho=0.7, trials=10, scale=None,
                       sample=True):
    """Besag (rook lattice) x AR1 group model, group-major (time outer), binomial(trials)
    logit, one observation per cell, 20 sum-to-zero rows (one per time point):
    Q = tau (R_ar1(rho) x Q*_besag), Q* the scaled Besag structure. With sample=False
    the counts stay zero (precision only)."""
    adj = rook_adjacency(nr, nr)
    R = besag_structure(adj)
    if scale is None:
        scale = float(scale_factors([R])[0])
    Qs = sp.csc_matrix(R * scale)
    I, M, O = ar1_pieces(T)
    mats = [sp.kron(I, Qs, format="csc"), sp.kron(M, Qs, format="csc"), sp.kron(O, Qs, format="csc")]
    indptr, indices = union_pattern(mats)
    comps = np.stack([values_on_pattern(m_, indptr, indices) for m_ in mats])
    n = nr * nr * T
    g = AffineGmrf(n, indptr, indices, comps, ["I x Q*", "M x Q*", "O x Q*"])
    rng = np.random.default_rng(seed)
    th_rho = math.log((1 + rho) / (1 - rho))  # inverse of rho_of
    c_true = tau * np.array([1.0, rho * rho, -rho])
    ns = nr * nr
    ca = np.zeros((T, n))
    for t in range(T):
        ca[t, t * ns:(t + 1) * ns] = 1.0
    y = np.zeros(n)
    if sample:
        # x ~ N(0, Q^-1) restricted to the constraints (sample, then project)
        x = sample_gmrf(sp.csc_matrix(g.csc(c_true) + sp.identity(n) * 1e-6), rng)
        x = x - ca.T @ np.linalg.solve(ca @ ca.T, ca @ x)
        y = rng.binomial(trials, 1.0 / (1.0 + np.exp(-x))).astype(float)

    def params(thetas):
        B = len(thetas)
        coefs = np.zeros((B, 3))
        dco = np.zeros((B, 2, 3))
        for b, th in enumerate(thetas):
            t_ = math.exp(th[0])
            r_, dr = rho_of(th[1])
            coefs[b] = t_ * np.array([1.0, r_ * r_, -r_])
            dco[b, 0] = coefs[b]
            dco[b, 1] = t_ * np.array([0.0, 2.0 * r_, -1.0]) * dr
        return coefs, dco, None, None

    lam_rho = pc_cor1_lambda(0.5, 0.75)

    def log_prior(th):
        # the reference's ar1 defaults (priors.rs:369-372): pc.prec(1, 0.01), pc.cor1(0.5, 0.75)
        a, ga = pc_prec_prior(th[0])
        b, gb = pc_cor1_prior(th[1], lam_rho)
        return a + b, np.array([ga, gb])

    return KronWorkload(g, None, y, np.full(n, float(trials)), None, "binomial", (ca, np.zeros(T)),
                        np.array([math.log(tau), th_rho]), params, log_prior)


def _space_time_projector(nx, T, n_obs, rng):
    """Barycentric projector of n_obs uniform points of uniform time slices of the
    nx x nx lattice (group-major latent index t nx^2 + s): (A (n_obs x nx^2 T, CSC),
    points, time slots)."""
    ns = nx * nx
    tslot = rng.integers(0, T, size=n_obs)
    pts = rng.uniform(size=(n_obs, 2))
    As = sp.csr_matrix(lattice_projector(nx, nx, pts))
    cnt = np.diff(As.indptr)
    cols = As.indices.astype(np.int64) + np.repeat(tslot.astype(np.int64) * ns, cnt)
    A = sp.csc_matrix(sp.csr_matrix((As.data, cols, As.indptr), shape=(n_obs, ns * T)))
    A.sort_indices()
    return A, pts, tslot


def spde_rw2_workload(nx=245, T=24, n_obs=500_000, seed=20241207, size=5.0, eps=1e-3, field="sample"):
    """SPDE alpha=2 (nx x nx lattice) x RW2 over T time steps, group-major; negative
    binomial (size) observations uniform in space x time through the barycentric
    projector of their time slice: Q = Q_spde(tau, kappa) x (R_rw2 + eps I).
    field="sample": eta = 0.5 + A x with x drawn from Q(theta_true) (test sizes);
    field="smooth": eta = 0.5 + 0.6 sin(2 pi s1) cos(2 pi s2) + 0.3 cos(2 pi t / T), no
    draw (the full size, where a host factorisation for the draw is out of reach)."""
    g_s, _, _ = spde_components(nx, nx)
    Rt = sp.csc_matrix(rw2_structure(T) + eps * sp.identity(T))
    mats = [sp.kron(Rt, sp.csc_matrix((g_s.components[k], g_s.indices, g_s.indptr), shape=(g_s.n, g_s.n)),
                    format="csc") for k in range(3)]
    indptr, indices = union_pattern(mats)
    comps = np.stack([values_on_pattern(m_, indptr, indices) for m_ in mats])
    del mats
    n = g_s.n * T
    g = AffineGmrf(n, indptr, indices, comps, ["Rt x C0", "Rt x G1", "Rt x G1C0invG1"])
    rng = np.random.default_rng(seed)
    # the pattern must contain A^T A (neighbours within a triangle of one time slice: already in G1)
    A, pts, tslot = _space_time_projector(nx, T, n_obs, rng)
    th = np.array([math.log(0.5), math.log(3.0)])
    if field == "sample":
        c, _ = spde_coefs(th[0], th[1])
        x = sample_gmrf(g.csc(c), rng)
        mu = np.exp(0.5 + A @ x)
    else:
        mu = np.exp(0.5 + 0.6 * np.sin(2 * math.pi * pts[:, 0]) * np.cos(2 * math.pi * pts[:, 1])
                    + 0.3 * np.cos(2 * math.pi * tslot / T))
    y = rng.negative_binomial(size, size / (size + mu)).astype(float)

    def params(thetas):
        B = len(thetas)
        coefs = np.zeros((B, 3))
        dco = np.zeros((B, 2, 3))
        for b, t_ in enumerate(thetas):
            cc, dc = spde_coefs(t_[0], t_[1])
            coefs[b] = cc
            dco[b] = dc
        return coefs, dco, None, None

    return KronWorkload(g, A, y, np.full(n_obs, math.exp(0.5)), np.full(n_obs, size), "nbinomial", None, th, params)


def spde_rw2_five_workload(nx=245, T=24, n_obs=500_000, seed=20241208, size=5.0, eps=1e-3, field="sample"):
    """The SPDE x RW2 space-time model with five hyperparameters, on the pattern of
    spde_rw2_workload (group-major, n = nx^2 T):

        Q(theta) = Q_spde(tau1, kappa1) x (R_rw2 + eps I) + Q_spde(tau2, kappa2) x I_T + tau3 I,

    a space-time smoothness term, a per-slice spatial term and a nugget;
    theta = (log tau1, log kappa1, log tau2, log kappa2, log tau3), seven affine
    components (the three SPDE matrices under each temporal factor, then I). Negative
    binomial (size) counts uniform in space x time through the barycentric projector of
    their slice; field="sample" draws x from Q(theta_true) (test sizes), "smooth" uses
    the deterministic field of spde_rw2_workload (full size). Priors: independent
    N(theta_true_k, 1)."""
    g_s, _, _ = spde_components(nx, nx)
    Rt = sp.csc_matrix(rw2_structure(T) + eps * sp.identity(T))
    It = sp.identity(T, format="csc")
    Cs = [sp.csc_matrix((g_s.components[k], g_s.indices, g_s.indptr), shape=(g_s.n, g_s.n)) for k in range(3)]
    mats = [sp.kron(Rt, Cs[k], format="csc") for k in range(3)] + \
           [sp.kron(It, Cs[k], format="csc") for k in range(3)] + [sp.identity(g_s.n * T, format="csc")]
    indptr, indices = union_pattern(mats)
    comps = np.stack([values_on_pattern(m_, indptr, indices) for m_ in mats])
    del mats
    n = g_s.n * T
    g = AffineGmrf(n, indptr, indices, comps, ["Rt x C0", "Rt x G1", "Rt x G1C0invG1",
                                               "It x C0", "It x G1", "It x G1C0invG1", "I"])
    rng = np.random.default_rng(seed)
    A, pts, tslot = _space_time_projector(nx, T, n_obs, rng)
    th = np.array([math.log(0.5), math.log(3.0), math.log(0.5), math.log(6.0), math.log(1.0)])

    def params(thetas):
        B = len(thetas)
        coefs = np.zeros((B, 7))
        dco = np.zeros((B, 5, 7))
        for b, t_ in enumerate(thetas):
            c1, d1 = spde_coefs(t_[0], t_[1])
            c2, d2 = spde_coefs(t_[2], t_[3])
            coefs[b, :3], coefs[b, 3:6], coefs[b, 6] = c1, c2, math.exp(t_[4])
            dco[b, 0:2, 0:3] = d1
            dco[b, 2:4, 3:6] = d2
            dco[b, 4, 6] = math.exp(t_[4])
        return coefs, dco, None, None

    def log_prior(t_):
        d = np.asarray(t_, dtype=float) - th
        return -0.5 * float(d @ d) - 2.5 * math.log(2 * math.pi), -d

    if field == "sample":
        c, _, _, _ = params(th[None])
        x = sample_gmrf(g.csc(c[0]), rng)
        mu = np.exp(0.5 + A @ x)
    else:
        mu = np.exp(0.5 + 0.6 * np.sin(2 * math.pi * pts[:, 0]) * np.cos(2 * math.pi * pts[:, 1])
                    + 0.3 * np.cos(2 * math.pi * tslot / T))
    y = rng.negative_binomial(size, size / (size + mu)).astype(float)
    return KronWorkload(g, A, y, np.full(n_obs, math.exp(0.5)), np.full(n_obs, size), "nbinomial", None, th, params,
                        log_prior)


def spde_rw2_posterior_precision(nx=245, T=24, n_obs=500_000, seed=20241207, size=5.0, eps=1e-3,
                                 log_tau=math.log(0.5), log_kappa=math.log(3.0), intercept=0.5):
    """Q_post = Q_spde(tau, kappa) x (R_rw2 + eps I) + A^T W A of the SPDE x RW2 negative
    binomial model at full size (n = nx^2 T), with the Fisher weights W = mu size / (size + mu)
    at eta = intercept for every observation; A is the barycentric projector of uniform
    points of a uniform time slice, built vectorised (no sampling of the field)."""
    g_s, _, _ = spde_components(nx, nx)
    c, _ = spde_coefs(log_tau, log_kappa)
    Qs = g_s.csc(c)
    Rt = sp.csc_matrix(rw2_structure(T) + eps * sp.identity(T))
    rng = np.random.default_rng(seed)
    tslot = rng.integers(0, T, size=n_obs)
    As = sp.csr_matrix(lattice_projector(nx, nx, rng.uniform(size=(n_obs, 2))))
    cnt = np.diff(As.indptr)
    cols = As.indices.astype(np.int64) + np.repeat(tslot.astype(np.int64) * g_s.n, cnt)
    n = g_s.n * T
    A = sp.csr_matrix((As.data, cols, As.indptr), shape=(n_obs, n))
    mu = math.exp(intercept)
    w = mu * size / (size + mu)
    Q = sp.csc_matrix(sp.kron(Rt, Qs, format="csc") + w * (A.T @ A))
    Q.sum_duplicates()
    Q.sort_indices()
    return Q


def scale_factor_dense(R) -> float:
    """scale_model factor of crates/inla_math/src/design.rs:151-195 through a dense
    eigen-decomposition (the reference's own route; small structures only):
    exp(mean log diag(R^+)), eigenvalues below sqrt(eps) max(lambda_max, 1) dropped."""
    lam, V = np.linalg.eigh(sp.csc_matrix(R).toarray())
    tol = math.sqrt(np.finfo(float).eps) * max(abs(max(lam.max(), 0.0)), 1.0)
    keep = lam > tol
    d = (V[:, keep] ** 2) @ (1.0 / lam[keep])
    d = d[(d > 0) & np.isfinite(d)]
    return float(math.exp(np.mean(np.log(d))))


KAPPA_HARD = 3.059023205018258e6  # HARD_CONSTRAINT_KAPPA, constraints.rs:268


@dataclass
class BesagPoissonWorkload:
    """Areal Poisson model: latent [intercept, covariate coefficient, u (regions)],
    u a scaled Besag ICAR with a sum-to-zero constraint, theta = log tau.

    ``gmrf`` components on the full pattern: 0 the fixed effects' precision
    (1e-4 on the two coefficients), 1 the scaled Besag structure Q* on u,
    2 the augmentation kappa a a^T of the constraint row a (the reference's explicit
    form, constraints.rs:275-313). ``gmrf_lowrank`` holds components 0 and 1 only, on
    their own sparse pattern, for the low-rank form of the augmentation."""

    gmrf: AffineGmrf
    gmrf_lowrank: AffineGmrf
    A: sp.csc_matrix
    y: np.ndarray
    x_cov: np.ndarray
    constraints: tuple
    scale: float
    theta_true: np.ndarray

    @property
    def n(self):
        return self.gmrf.n

    def params(self, thetas, lowrank=False):
        thetas = np.atleast_2d(np.asarray(thetas, dtype=float))
        B = len(thetas)
        K = 2 if lowrank else 3
        coefs = np.zeros((B, K))
        dco = np.zeros((B, 1, K))
        for b, th in enumerate(thetas):
            t = math.exp(th[0])
            coefs[b, 0] = 1.0
            coefs[b, 1] = t
            if not lowrank:
                coefs[b, 2] = 1.0
            dco[b, 0, 1] = t
        return coefs, dco, None, None

    def params_lowrank(self, thetas):
        return self.params(thetas, lowrank=True)

    def log_prior(self, theta):
        # the reference's default precision prior loggamma(1, 5e-5) (api.py:1357)
        v, g = loggamma_prior(float(np.asarray(theta).ravel()[0]), 1.0, 5e-5)
        return v, np.array([g])

    def prior_without_kappa(self, theta):
        """Q(theta) without the augmentation (the oracle adds kappa a a^T itself)."""
        c = self.params(np.atleast_1d(theta))[0][0]
        c[2] = 0.0
        return self.gmrf.csc(c)


def besag_poisson_workload(nr=30, seed=20241208, intercept=0.5, beta=0.3, tau=1.0,
                           fixed_prec=1e-4):
    """Besag ICAR + Poisson on an nr x nr rook lattice (node = r nr + c), scale_model,
    sum-to-zero on u, intercept 0.5 and a N(0, 1) covariate with coefficient 0.3,
    u drawn from the scaled ICAR with tau = 1, y ~ Poisson(exp(0.5 + 0.3 x + u)).
    Latent layout as the reference's (fixed effects first, api.py:1508-1537)."""
    adj = rook_adjacency(nr, nr)
    R = besag_structure(adj)
    ns = nr * nr
    n = ns + 2
    fac = scale_factor_dense(R)
    Qs = sp.csc_matrix(R * fac)
    rng = np.random.default_rng(seed)
    x_cov = rng.standard_normal(ns)
    # u ~ N(0, (tau Q*)^+): the eigenvectors of the non-zero eigenvalues
    lam, V = np.linalg.eigh(Qs.toarray())
    keep = lam > 1e-9 * lam.max()
    u = V[:, keep] @ (rng.standard_normal(int(keep.sum())) / np.sqrt(tau * lam[keep]))
    y = rng.poisson(np.exp(intercept + beta * x_cov + u)).astype(float)
    # A = [1, x, I]: row i is region i
    rows = np.r_[np.arange(ns), np.arange(ns), np.arange(ns)]
    cols = np.r_[np.zeros(ns, dtype=int), np.ones(ns, dtype=int), 2 + np.arange(ns)]
    vals = np.r_[np.ones(ns), x_cov, np.ones(ns)]
    A = sp.csc_matrix((vals, (rows, cols)), shape=(ns, n))
    A.sort_indices()
    Kfix = sp.csc_matrix(sp.diags(np.r_[fixed_prec, fixed_prec, np.zeros(ns)]))
    Kb = sp.csc_matrix(sp.block_diag([sp.csc_matrix((2, 2)), Qs]))
    a = np.zeros((1, n))
    a[0, 2:] = 1.0 / math.sqrt(ns)  # sum_to_zero_constraint (constraints.rs:113-127)
    Kk = sp.csc_matrix(KAPPA_HARD * (a.T @ a))
    ata = sp.csc_matrix(A.T @ A)
    indptr, indices = union_pattern([Kfix, Kb, Kk, ata])
    comps = np.stack([values_on_pattern(m, indptr, indices) for m in (Kfix, Kb, Kk)])
    g = AffineGmrf(n, indptr, indices, comps, ["fixed", "Q* (u)", "kappa a a^T"])
    ip2, ix2 = union_pattern([Kfix, Kb, ata])
    g2 = AffineGmrf(n, ip2, ix2, np.stack([values_on_pattern(m, ip2, ix2) for m in (Kfix, Kb)]),
                    ["fixed", "Q* (u)"])
    return BesagPoissonWorkload(g, g2, A, y, x_cov, (a, np.zeros(1)), fac, np.array([math.log(tau)]))
