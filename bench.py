#!/usr/bin/env python
"""Headline benchmark: CCD theta-points/s (LDL^T + selected inversion + gradient), FP64.

Workload: SPDE alpha=2 on a 142 x 142 unit-square lattice (n = 20,164), 50,000
uniform observations through the barycentric projector, Gaussian likelihood
(sigma_eps = 0.1), true range 0.3 and sigma 1, field sampled from Q(tau, kappa);
internal theta = (log prec_eps, log tau, log kappa); CCD with m = 3 gives 15
theta-points around the mode, each evaluated as one unit of work:
  prior LDL^T (Q + 1e-12 I), posterior LDL^T (Q + A^T W A), the mode solve,
  posterior and prior selected inversion, the gradient of the log marginal
  likelihood (trace contractions), marginal variances.
One step = all 15 theta-points in one batched call; value = theta-points / s.

The CCD design is scaled with the Hessian from the finite-difference stencil of
crates/inla_math/src/hyper_opt.rs:197-259, all 19 stencil points evaluated in one
batch before timing. The mode is the generating theta (the design does not
depend on how the mode was found).
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-inla_amd"))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X dense FP64 (MFMA), FLOP/s
PEAK_HBM = 8.0e12     # MI355X HBM3E, B/s
METRIC = "CCD θ-points/sec (LDLᵀ+selinv+grad) FP64, SPDE n≈20k; % FP64-MFMA/HBM peak"
# GMRF_ORDER_* of the ordering option; the analysis' order_kind names
ORDERS = {"chooser": 0, "amd": 1, "nd": 4, "auto": 5}
ORDER_KINDS = {0: "minimum degree", 1: "time-major", 2: "RCM", 3: "natural", 4: "given", 5: "nested dissection"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_cores():
    """Host cores this process may use (the affinity mask), at most 16: the GPU box's
    CPU share per GPU."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(wl, thetas):
    """The CPU restatement of the reference (the oracle) on the same CCD theta-points,
    the reference's way: each node is an independent find_mode + Takahashi diagonal
    (inference.rs:1306-1385), run across the host cores as its Rayon par_iter over
    nodes does (the oracle's calls release the interpreter lock, so a thread pool
    runs them in parallel), and the gradient by central differences, 2m further
    Laplace evaluations per point (hyper_opt.rs:197-259). As the reference's
    factorize_sparse (sparse_ldlt.rs:200-207,234), the oracle keeps the ordering of the
    last pattern factorised on each thread and reuses it when the pattern repeats, so
    each thread orders the pattern once. Three measurements: node work + FD gradient
    on all cores for all points (the value, the unit the GPU line reports), node work
    alone on all cores, and node work on one core for one point with the ordering
    already cached (the first point's time, ordering included, is reported beside
    it)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from oracle import oracle as orc

    cores = _cpu_cores()
    m = thetas.shape[1]
    h = 1e-4

    def node(th):
        c, _, lp, _ = wl.params(th[None])
        r = orc.find_mode(wl.gmrf.csc(c[0]), "gaussian", "identity", wl.y, lp[0], A=wl.A, keep_post=True)
        r.post_factor().takahashi_diag()

    def fd_eval(th):
        c, _, lp, _ = wl.params(th[None])
        orc.find_mode(wl.gmrf.csc(c[0]), "gaussian", "identity", wl.y, lp[0], A=wl.A)

    fd_pts = []
    for th in thetas:
        for l in range(m):
            for sgn in (h, -h):
                t = np.array(th, dtype=float)
                t[l] += sgn
                fd_pts.append(t)
    t0 = time.perf_counter()
    node(thetas[0])
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    node(thetas[1])
    t_one = time.perf_counter() - t0
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(node, thetas))
        t_nodes = time.perf_counter() - t0
        t0 = time.perf_counter()
        list(ex.map(fd_eval, fd_pts))
        t_fd = time.perf_counter() - t0
    npts = len(thetas)
    return {"value": npts / (t_nodes + t_fd), "unit": "theta-points/s", "cores": cores, "kind": "port",
            "sample": f"the {npts} CCD theta-points of the same workload on the oracle, a thread per node "
                      f"over {cores} host cores: node work (prior + posterior LDLT, mode, Takahashi diag) "
                      f"{t_nodes:.1f} s + 2m = {2 * m} finite-difference Laplace evaluations per point for the "
                      f"gradient {t_fd:.1f} s",
            "node_only": {"value": npts / t_nodes, "unit": "theta-points/s", "cores": cores},
            "one_core_node_only": {"value": 1.0 / t_one, "unit": "theta-points/s", "cores": 1,
                                   "sample": f"1 theta-point with the ordering cached, {t_one:.2f} s "
                                             f"(the first point, ordering included: {t_first:.2f} s)"},
            "host_cpus_visible": os.cpu_count()}


class OracleObjective:
    """The interface of inla_gpu.fit.Objective on the oracle: the log posterior
    (mlik + log prior) of the Gaussian SPDE workload, the points of one call spread
    over the host cores (one thread per point, as the reference's par_iter over CCD
    nodes)."""

    def __init__(self, wl, cores):
        self.wl, self.cores = wl, cores
        self.calls = 0
        self.points = 0

    def values(self, thetas, grad=False, want_x=False, want_var=False):
        from concurrent.futures import ThreadPoolExecutor

        import numpy as np

        from oracle import oracle as orc

        thetas = np.atleast_2d(np.asarray(thetas, dtype=float))
        wl = self.wl

        def one(th):
            c, _, lp, _ = wl.params(th[None])
            r = orc.find_mode(wl.gmrf.csc(c[0]), "gaussian", "identity", wl.y, lp[0], A=wl.A, keep_post=want_var)
            var = r.post_factor().takahashi_diag() if want_var else None
            return r.mlik + wl.log_prior(th)[0], r.x, var

        with ThreadPoolExecutor(min(self.cores, len(thetas))) as ex:
            out = list(ex.map(one, thetas))
        self.calls += 1
        self.points += len(thetas)
        if self.calls % 10 == 0:
            log(f"[cpu fit] {self.calls} calls, {self.points} theta-points")
        res = {"g": np.array([o[0] for o in out])}
        if want_x:
            res["x"] = np.array([o[1] for o in out])
        if want_var:
            res["var"] = np.array([o[2] for o in out])
        return res


def cpu_fit_line(wl, theta0):
    """The reference's fit sequence on the oracle, measured: Nelder-Mead one point at
    a time (hyper_opt.rs:15-185), the Hessian stencil point by point (hyper_opt.rs:197-259),
    then every CCD node at once across the host cores (inference.rs:1394) and the
    mixture."""
    from inla_gpu import ccd
    from inla_gpu import fit as F

    cores = _cpu_cores()
    obj = OracleObjective(wl, cores)
    t0 = time.perf_counter()
    mode = F.nelder_mead_mode(obj, theta0)
    t1 = time.perf_counter()
    log(f"[cpu fit] Nelder-Mead: {t1 - t0:.1f} s, {obj.calls} evaluations")
    h = F._hessian_step(mode)
    import numpy as np

    pts = ccd.hessian_points(mode, h)
    H = ccd.hessian_from_values(np.concatenate([obj.values(p[None])["g"] for p in pts]), len(mode), h)
    t2 = time.perf_counter()
    thetas, w = F._nodes(mode, H)
    r = obj.values(thetas, want_x=True, want_var=True)
    _, mean, var, mlik = ccd.aggregate(w, r["g"], r["x"], r["var"])
    t3 = time.perf_counter()
    return {"total_s": t3 - t0, "mode_s": t1 - t0, "hessian_s": t2 - t1, "nodes_s": t3 - t2,
            "evaluations": obj.points, "cores": cores, "mode": [float(v) for v in mode], "mlik": float(mlik)}


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "pmc_traffic_spde_gaussian_n20164.md")


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None when it is not listed."""
    try:
        with open(TRAFFIC_TABLE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 5 and cells[0] == kernel:
                    return float(cells[4])
    except OSError:
        return None
    return None


def plain_gpu(model, wl, thetas, reps=3):
    """The plain per-point GPU path on the same theta-points: one point per call
    (prior + posterior LDL^T, mode, selected-inverse diagonal), and the gradient by
    central differences (2m further calls), as a one-theta-at-a-time solver does."""
    import numpy as np
    import torch

    h = 1e-4
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for th in thetas[:reps]:
        c, _, lp, _ = wl.params(th[None])
        model.eval(c, lpar=lp, want_x=True, want_var=True)
        for l in range(len(th)):
            for s in (h, -h):
                tt = np.array(th, dtype=float)
                tt[l] += s
                c, _, lp, _ = wl.params(tt[None])
                model.eval(c, lpar=lp, want_x=False, want_var=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": reps / dt, "unit": "theta-points/s", "ms_per_point": dt * 1e3 / reps,
            "sample": f"{reps} CCD theta-points one at a time: node work + {2 * len(thetas[0])} "
                      f"finite-difference evaluations each"}


def fit_compare(model, wl):
    """End-to-end fit (mode search + Hessian + all CCD nodes + mixture) from the same
    start: batched BFGS on the adjoint gradient vs the reference's one-point-at-a-time
    Nelder-Mead sequence."""
    import numpy as np

    from inla_gpu import fit as F

    theta0 = wl.theta_true + np.array([0.5, -0.3, 0.3])
    ob = F.Objective(model, wl.params, wl.log_prior)
    rb = F.fit_batched(ob, theta0)
    op = F.Objective(model, wl.params, wl.log_prior)
    rp = F.fit_plain(op, theta0)
    out = {"batched_s": rb.seconds["total"], "plain_s": rp.seconds["total"],
           "speedup": rp.seconds["total"] / rb.seconds["total"],
           "batched_calls": rb.n_evals, "batched_points": rb.n_points,
           "plain_calls": rp.n_evals, "plain_points": rp.n_points,
           "mode_batched": [float(v) for v in rb.mode], "mode_plain": [float(v) for v in rp.mode],
           "mlik_batched": float(rb.mlik), "mlik_plain": float(rp.mlik),
           "start": [float(v) for v in theta0]}
    return out, rb


def poisson_line(nx, n_obs, steps=5, order="auto"):
    """The batched Poisson configuration on the same mesh (intercept + SPDE field,
    y ~ Poisson(exp(1 + u)), theta = (log tau, log kappa), CCD m = 2: 9 points):
    every point's prior LDL^T, Newton iterates, selected inversions and gradient
    in one batched call."""
    import numpy as np
    import torch

    from inla_gpu import ccd
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload(nx=nx, n_obs=n_obs)
    g = wl.gmrf
    m = len(wl.theta_true)
    z, _ = ccd.ccd_design(m, 1.0)
    hh = max(1e-4, 0.05 * float(np.max(np.abs(wl.theta_true))))
    stencil = ccd.hessian_points(wl.theta_true, hh)
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A,
                         capacity=max(len(z), len(stencil)), order=order)
    c, _, _, _ = wl.params(stencil)
    r = model.eval(c, want_x=False, want_var=False)
    lpost = r["mlik"].cpu().numpy() + np.array([wl.log_prior(t)[0] for t in stencil])
    lam, V = ccd.scaling_from_hessian(ccd.hessian_from_values(lpost, m, hh))
    thetas = np.array([ccd.theta_from_z(wl.theta_true, lam, V, zz) for zz in z])
    coefs, dcoefs, _, _ = wl.params(thetas)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    coefs, dcoefs = dev(coefs), dev(dcoefs)
    out = model.eval(coefs, dcoefs=dcoefs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = model.eval(coefs, dcoefs=dcoefs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"value": len(thetas) / dt, "unit": "theta-points/s", "ms_per_step": dt * 1e3,
            "theta_points": len(thetas), "n": g.n, "nnz_q": int(len(g.indices)),
            "newton_iters": [int(v) for v in out["n_iter"]], "status_ok": bool((out["status"] == 0).all()),
            "workload": f"spde-poisson lattice {nx}x{nx} + intercept (n={g.n}), {n_obs} obs, CCD m=2"}


def hetero_compare(k=64, reps=3):
    """Heterogeneous batch of k BYM2 + Poisson precision matrices (Delaunay-neighbour
    graphs, sizes uniform in [500, 20000], phi = 0.5, Poisson weights at u = 0):
    factorise + selected-inverse diagonal + one solve for all k, as one multi-pattern
    launch sequence vs k separate single-pattern batches."""
    import numpy as np
    import scipy.sparse as sp
    import torch

    from inla_gpu import workloads as W
    from inla_gpu.solver import Batch, MultiBatch, Symbolic

    models = W.bym2_hetero_batch(k=k, sample=False)
    mats = []
    for mdl in models:
        Q = sp.csc_matrix((mdl.values(1.0, 0.5), mdl.indices, mdl.indptr), shape=(mdl.n, mdl.n))
        Q = sp.csc_matrix(Q + math.e * sp.identity(mdl.n))
        Q.sort_indices()
        mats.append(Q)
    mb = MultiBatch(mats, capacity=1)
    vals = torch.as_tensor(np.concatenate([q.data for q in mats])[None], device="cuda")
    rhs = torch.ones((1, mb.n), dtype=torch.float64, device="cuda")

    def run_multi():
        mb.factorize(vals)
        mb.solve(rhs)
        mb.selinv()
        return mb.selinv_diag()

    singles = [(Batch(Symbolic.from_csc(q), 1), torch.as_tensor(q.data[None], device="cuda"),
                torch.ones((1, q.shape[0]), dtype=torch.float64, device="cuda")) for q in mats]

    def run_single():
        for bt, v, r in singles:
            bt.factorize(v)
            bt.solve(r)
            bt.selinv()
            bt.selinv_diag()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    tm, ts = timed(run_multi), timed(run_single)
    ld, st = mb.logdet()
    return {"k": k, "n_total": mb.n, "nnz_total": mb.nnz, "sizes": [int(q.shape[0]) for q in mats[:8]],
            "multi_ms": tm * 1e3, "separate_ms": ts * 1e3, "speedup": ts / tm,
            "models_per_s": k / tm, "status_ok": bool(np.all(st == 0)),
            "work": "LDL^T + selected-inverse diagonal + one solve per model"}


_HETERO = {}
HETERO_ORDER = ["chooser"]  # ordering of the heterogeneous models (set from --hetero-order)


def _hetero_setup(k):
    """The 64-model BYM2 + Poisson batch, its union model (capacity 9) and the models
    alone, built once per process."""
    if k in _HETERO:
        return _HETERO[k]
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    t0 = time.perf_counter()
    wl = W.bym2_poisson_batch(k=k)
    t_build = time.perf_counter() - t0
    log(f"   workload: {k} models, n = {wl.n}, {t_build:.1f} s")
    t0 = time.perf_counter()
    union = LaplaceModel(wl.n, wl.indptr, wl.indices, wl.components, wl.y, family="poisson", A=wl.A,
                         blocks=wl.block_ptr, capacity=9, order=HETERO_ORDER[0])
    t_setup = time.perf_counter() - t0
    bp = wl.block_ptr
    singles = []
    for q in range(k):
        e0, e1 = wl.indptr[bp[q]], wl.indptr[bp[q + 1]]
        singles.append(LaplaceModel(bp[q + 1] - bp[q], wl.indptr[bp[q]:bp[q + 1] + 1] - e0,
                                    wl.indices[e0:e1] - bp[q], wl.components[:, e0:e1], wl.block_y(q),
                                    family="poisson", A=wl.block_A(q), capacity=9, order=HETERO_ORDER[0]))
    _HETERO[k] = (wl, union, singles, t_build, t_setup)
    return _HETERO[k]


def hetero_laplace_line(k=64, reps=3, cpu_sample=True):
    """The heterogeneous Laplace batch: k BYM2 + Poisson models with an intercept each
    (Delaunay-neighbour graphs, sizes uniform in [500, 20000], phi = 0.5, seeded), 9 CCD
    theta-points per model (m = 2, theta = 0.5 z around (log tau, logit phi) = (0, 0)):
    Newton, mlik, marginal variances and the adjoint gradient of every (point, model)
    pair as one block-diagonal model, against the k models evaluated one after another
    on the same GPU, and (cpu_sample) the oracle on the models' centre points across the
    host cores (node work: Newton, mlik, Takahashi diagonal; no gradient)."""
    import numpy as np
    import torch

    from inla_gpu import ccd

    wl, union, singles, t_build, t_setup = _hetero_setup(k)
    z, _ = ccd.ccd_design(2, 1.0)
    B = len(z)
    thetas = np.broadcast_to((0.5 * z)[:, None, :], (B, k, 2)).copy()
    c, dc, _, _ = wl.params(thetas)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    cd, dcd = dev(c), dev(dc)

    def run_union():
        return union.eval(cd, dcoefs=dcd)

    cq = [dev(c[:, q]) for q in range(k)]
    dcq = [dev(dc[:, q]) for q in range(k)]

    def run_single():
        return [singles[q].eval(cq[q], dcoefs=dcq[q]) for q in range(k)]

    def timed(fn):
        out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    tu, ru = timed(run_union)
    log(f"   union: {tu * 1e3:.1f} ms per evaluation of {B} x {k} (point, model) pairs")
    ts, rs = timed(run_single)
    log(f"   separate models: {ts * 1e3:.1f} ms")
    ml_u = ru["mlik"].cpu().numpy()
    ml_s = np.stack([r["mlik"].cpu().numpy() for r in rs], axis=1)
    npts = B * k
    res = {"value": npts / tu, "unit": "theta-points/s", "ms_per_step": tu * 1e3,
           "theta_points": npts, "models": k, "points_per_model": B, "n_total": int(wl.n),
           "nnz_q_total": int(len(wl.indices)), "sizes_min_max": [int(min(m.n for m in wl.models)),
                                                                   int(max(m.n for m in wl.models))],
           "newton_iters_min_max": [int(ru["n_iter"].min()), int(ru["n_iter"].max())],
           "status_ok": bool((ru["status"] == 0).all()),
           "separate_models": {"value": npts / ts, "ms_per_step": ts * 1e3},
           "speedup_vs_separate": ts / tu,
           "max_abs_mlik_diff_union_vs_separate": float(np.max(np.abs(ml_u - ml_s))),
           "setup_s": {"workload": t_build, "model": t_setup},
           "work": "per (theta-point, model): Newton to the mode, mlik, marginal variances, adjoint gradient",
           "workload": f"{k} BYM2 + Poisson models (intercept each), Delaunay graphs, sizes U[500, 20000], "
                       f"CCD m=2 ({B} theta-points per model)"}
    if cpu_sample:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle as orc

        cores = _cpu_cores()

        def node(q):
            r = orc.find_mode(wl.block_prior(q, thetas[0, q]), "poisson", "log", wl.block_y(q), 1.0,
                              A=wl.block_A(q), keep_post=True)
            r.post_factor().takahashi_diag()
            return r.mlik

        with ThreadPoolExecutor(cores) as ex:
            t0 = time.perf_counter()
            ml_o = np.array(list(ex.map(node, range(k))))
            t_cpu = time.perf_counter() - t0
        res["cpu_sample"] = {"value": k / t_cpu, "unit": "theta-points/s", "cores": cores, "kind": "port",
                             "sample": f"the {k} models' centre theta-points on the oracle, a thread per model "
                                       f"over {cores} host cores, node work only (no gradient): {t_cpu:.1f} s",
                             "max_abs_mlik_diff_vs_gpu": float(np.max(np.abs(ml_o - ml_u[0])))}
    return res


def hetero_fit_line(k=64):
    """The 64 BYM2 + Poisson models fitted at once (fit_blocks): each model's BFGS mode
    on the adjoint gradient, its Hessian stencil and its 9 CCD nodes with variances, every
    step one block evaluation of the union; against each model fitted alone with the same
    driver (fit_batched on its own model, one model after another)."""
    import numpy as np
    import torch

    from inla_gpu import fit as F
    from inla_gpu import workloads as W

    wl, union, singles, _, _ = _hetero_setup(k)
    theta0 = np.tile([0.5, 0.5], (k, 1))
    obj = F.BlockObjective(union, wl.params, W.bym2_log_prior)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fits, secs = F.fit_blocks(obj, theta0, wl.block_ptr)
    torch.cuda.synchronize()
    t_union = time.perf_counter() - t0
    log(f"   all {k} fits as blocks: {t_union:.2f} s ({secs['bfgs_rounds']} BFGS rounds, {obj.calls} calls)")
    t0 = time.perf_counter()
    alone = []
    for q in range(k):
        def params_q(th, q=q):
            c, dc, _, _ = wl.params(np.repeat(np.atleast_2d(th)[:, None, :], k, axis=1))
            return c[:, q], dc[:, q], None, None

        alone.append(F.fit_batched(F.Objective(singles[q], params_q, W.bym2_log_prior), theta0[q]))
    torch.cuda.synchronize()
    t_alone = time.perf_counter() - t0
    log(f"   the {k} models fitted one after another: {t_alone:.2f} s")
    dmode = max(float(np.max(np.abs(f.mode - a.mode))) for f, a in zip(fits, alone))
    dmean = max(float(np.max(np.abs(f.mean - a.mean)) / np.max(np.abs(a.mean))) for f, a in zip(fits, alone))
    modes = np.array([f.mode for f in fits])
    return {"value": k / t_union, "unit": "models fitted/s", "seconds": t_union,
            "seconds_by_phase": {kk: secs[kk] for kk in ("mode", "hessian", "nodes")},
            "bfgs_rounds": secs["bfgs_rounds"], "block_evaluations": obj.calls,
            "models_alone_seconds": t_alone, "speedup_vs_alone": t_alone / t_union,
            "max_abs_mode_diff_vs_alone": dmode, "max_rel_mean_diff_vs_alone": dmean,
            "mode_range": [modes.min(axis=0).tolist(), modes.max(axis=0).tolist()],
            "work": "per model: BFGS on the adjoint gradient from (0.5, 0.5), Hessian stencil, 9 CCD nodes with "
                    "variances, mixture",
            "workload": f"{k} BYM2 + Poisson models (intercept each), Delaunay graphs, sizes U[500, 20000]; "
                        "priors loggamma(1, 5e-5) on tau, N(0, 1) on logit phi"}


def big_solver_line(Q, label, probes=6, order="auto"):
    """Full-size solver line at batch 1: factorise, solve and selected-inverse diagonal
    of one posterior precision, with size-independent checks: log|2Q| - log|Q| = n log 2,
    the solve residual, and diagonal entries against solves of e_i."""
    import numpy as np
    import torch

    from inla_gpu.solver import Batch, Symbolic

    n = Q.shape[0]
    t0 = time.perf_counter()
    sym = Symbolic.from_csc(Q, order=ORDERS[order])
    st = sym.stats()
    t_sym = time.perf_counter() - t0
    bt = Batch(sym, 1)
    t_setup = time.perf_counter() - t0
    print(f"[{label}] analysed in {t_sym:.1f} s, workspaces in {t_setup - t_sym:.1f} s", file=sys.stderr,
          flush=True)
    v = torch.as_tensor(Q.data[None], device="cuda")
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n)
    bd = torch.as_tensor(b[None], device="cuda")

    def timed(fn):
        torch.cuda.synchronize()
        a = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - a

    _, tf = timed(lambda: bt.factorize(v))
    ld1, st1 = bt.logdet()
    print(f"[{label}] factorised in {tf:.2f} s", file=sys.stderr, flush=True)
    x, ts = timed(lambda: bt.solve(bd))
    _, tsel = timed(lambda: bt.selinv())
    print(f"[{label}] selected inversion in {tsel:.2f} s", file=sys.stderr, flush=True)
    dg = bt.selinv_diag()[0].cpu().numpy()
    xr = x[0].cpu().numpy()
    resid = float(np.linalg.norm(Q @ xr - b) / np.linalg.norm(b))
    idx = rng.choice(n, probes, replace=False)
    dmax = 0.0
    for i in idx:
        e = torch.zeros((1, n), dtype=torch.float64, device="cuda")
        e[0, i] = 1.0
        dmax = max(dmax, abs(bt.solve(e)[0, i].item() - dg[i]) / dg[i])
    bt.factorize(2.0 * v)
    ld2, _ = bt.logdet()
    return {"workload": label, "n": int(n), "nnz_q": int(Q.nnz), "order": order,
            "order_kind": int(st["order_kind"]), "nsuper": int(st["nsuper"]), "levels": int(st["nlevels"]),
            "nnz_l": int(st["nnz_l"]), "l_size": int(st["l_size"]), "max_front": int(st["max_front"]),
            "flops_ldlt": st["flops_ldlt"], "flops_selinv": st["flops_selinv"],
            "analyse_s": t_sym, "setup_s": t_setup, "factorize_s": tf, "solve_s": ts, "selinv_s": tsel,
            "ldlt_tflops": st["flops_ldlt"] / tf / 1e12, "selinv_tflops": st["flops_selinv"] / tsel / 1e12,
            "status": int(st1[0]), "logdet": float(ld1[0]),
            "logdet_scaling_err": float(abs((ld2[0] - ld1[0]) - n * math.log(2.0))),
            "solve_rel_residual": resid, "selinv_diag_vs_solves_rel": dmax}


def c4_solver_line(nr=100, T=20, probes=6, order="auto"):
    """Full-size Besag x AR1 (100 x 100 rook lattice x T = 20, n = 200,000; the
    reference's chooser picks RCM: a chain of dense fronts): Q_post = Q(tau=2, rho=0.7)
    + diag(2.5) (binomial(10) weights at eta = 0)."""
    import scipy.sparse as sp

    from inla_gpu import workloads as W

    wl = W.besag_ar1_workload(nr=nr, T=T, sample=False)
    c, _, _, _ = wl.params(wl.theta_true[None])
    Q = sp.csc_matrix(wl.gmrf.csc(c[0]) + 2.5 * sp.identity(wl.n))
    Q.sort_indices()
    return big_solver_line(Q, f"besag {nr}x{nr} x ar1(T={T}), n={wl.n}, Q_post at tau=2, rho=0.7", probes, order)


def c4_fit_line(order="auto", nr=100, T=20):
    """Full-size Besag x AR1 binomial(10) (100 x 100 rook lattice x T = 20, n = 200,000) with
    its 20 sum-to-zero rows applied as the low-rank augmentation (the explicit
    kappa A_c^T A_c has 2e9 entries): one batched evaluation of the 9 CCD points with
    the adjoint gradient (theta-points/s, also under the reference chooser's RCM
    order), and the end-to-end fit (mode, Hessian, CCD nodes, mixture): batched BFGS
    on the adjoint gradient vs the one-point Nelder-Mead sequence. Priors: the
    reference's ar1 defaults pc.prec(1, 0.01) and pc.cor1(0.5, 0.75)."""
    import numpy as np
    import torch

    from inla_gpu import ccd
    from inla_gpu import fit as F
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    t0 = time.perf_counter()
    wl = W.besag_ar1_workload(nr=nr, T=T)
    ca, ce = wl.constraints
    ca = ca / math.sqrt(nr * nr)
    g = wl.gmrf
    log(f"[c4 fit] workload generated in {time.perf_counter() - t0:.1f} s")

    def make(order_):
        return LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="binomial", obs_a=wl.obs_a,
                            constraints=(ca, ce), constraint_kappa=3.059023205018258e6, capacity=9,
                            order=order_)

    model = make(order)
    m = len(wl.theta_true)
    hh = max(1e-4, 0.05 * float(np.max(np.abs(wl.theta_true))))
    stencil = ccd.hessian_points(wl.theta_true, hh)
    ob = F.Objective(model, wl.params, wl.log_prior)
    lpost = ob.values(stencil)["g"]
    lam, V = ccd.scaling_from_hessian(ccd.hessian_from_values(lpost, m, hh))
    z, _ = ccd.ccd_design(m, 1.0)
    thetas = np.array([ccd.theta_from_z(wl.theta_true, lam, V, zz) for zz in z])
    coefs, dcoefs, _, _ = wl.params(thetas)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    coefs, dcoefs = dev(coefs), dev(dcoefs)

    def timed_eval(mod):
        torch.cuda.synchronize()
        a = time.perf_counter()
        out = mod.eval(coefs, dcoefs=dcoefs)
        torch.cuda.synchronize()
        return out, time.perf_counter() - a

    timed_eval(model)
    out, dt = timed_eval(model)
    res = {"workload": f"besag {nr}x{nr} x ar1(T={T}) binomial(10), n={g.n}, 20 sum-to-zero rows (low rank)",
           "order": order, "theta_points": len(thetas), "eval_s": dt, "value": len(thetas) / dt,
           "unit": "theta-points/s", "newton_iters": [int(v) for v in out["n_iter"]],
           "status_ok": bool((out["status"] == 0).all())}
    log(f"[c4 fit] batched 9-point evaluation: {dt:.2f} s")
    del model
    if order != "chooser":
        mc = make("chooser")
        oc, dtc = timed_eval(mc)
        res["chooser_eval_s"] = dtc
        res["chooser_vs_order_mlik_maxdiff"] = float((oc["mlik"] - out["mlik"]).abs().max().item())
        log(f"[c4 fit] the same under the chooser's order: {dtc:.2f} s")
        del mc
        model = make(order)
    class Logged(F.Objective):  # a progress line every 10 calls (long one-point sequences)
        def values(self, *a, **k):
            if self.calls % 10 == 0:
                log(f"[c4 fit] {self.calls} calls, {self.points} theta-points")
            return super().values(*a, **k)

    theta0 = wl.theta_true + np.array([0.5, -0.3])
    rb = F.fit_batched(Logged(model, wl.params, wl.log_prior), theta0)
    log(f"[c4 fit] batched BFGS fit {rb.seconds['total']:.1f} s")
    rp = F.fit_plain(Logged(model, wl.params, wl.log_prior), theta0)
    log(f"[c4 fit] Nelder-Mead fit {rp.seconds['total']:.1f} s")
    res.update({"fit_batched_s": rb.seconds["total"], "fit_plain_s": rp.seconds["total"],
                "fit_speedup": rp.seconds["total"] / rb.seconds["total"],
                "batched_calls": rb.n_evals, "batched_points": rb.n_points,
                "plain_calls": rp.n_evals, "plain_points": rp.n_points,
                "mode_batched": [float(v) for v in rb.mode], "mode_plain": [float(v) for v in rp.mode],
                "mlik_batched": float(rb.mlik), "mlik_plain": float(rp.mlik),
                "theta_true": [float(v) for v in wl.theta_true]})
    return res


def c5_solver_line(nx=245, T=24, n_obs=500_000, probes=4, order="auto"):
    """Full-size SPDE x RW2 (245 x 245 lattice x 24 time steps, n = 1,440,600; RCM chain of
    fronts up to 13k wide, about 100 GB of L): Q_post with the negative binomial (size 5)
    Fisher weights of 500,000 observations at eta = 0.5."""
    from inla_gpu import workloads as W

    t0 = time.perf_counter()
    Q = W.spde_rw2_posterior_precision(nx=nx, T=T, n_obs=n_obs)
    print(f"[c5] Q_post built in {time.perf_counter() - t0:.1f} s (nnz {Q.nnz})", file=sys.stderr, flush=True)
    return big_solver_line(Q, f"spde {nx}x{nx} x rw2(T={T}), n={Q.shape[0]}, {n_obs} NB obs, "
                              "Q_post at tau=0.5, kappa=3", probes, order)


def c5_laplace_line(nx=245, T=24, n_obs=500_000, order="auto", fd_components=(1, 4)):
    """Full-size SPDE x RW2 negative binomial Laplace evaluation with five hyperparameters
    (n = 1,440,600, 500,000 observations, size 5; Q = Q_spde(tau1, kappa1) x (R_rw2 +
    eps I) + Q_spde(tau2, kappa2) x I_T + tau3 I) at one theta-point with the adjoint
    gradient and the marginal variances, in the lean form (prior factorised and inverted
    alone first: one factor and one selected inverse on the device at a time).
    Synthetic counts from a smooth space-time field (no draw from Q at this size). The
    gradient components in fd_components against central differences of further
    evaluations (two steps and their Richardson extrapolation). The m = 5 CCD design has
    27 nodes; one node per call is what the device holds at this size."""
    import numpy as np
    import torch

    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    t0 = time.perf_counter()
    wl = W.spde_rw2_five_workload(nx=nx, T=T, n_obs=n_obs, field="smooth")
    g = wl.gmrf
    log(f"[c5 laplace] workload n={g.n} nnz={len(g.indices)} K={g.components.shape[0]} in "
        f"{time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="nbinomial", A=wl.A,
                         obs_a=wl.obs_a, obs_b=wl.obs_b, capacity=1, lean=True, order=order)
    t_setup = time.perf_counter() - t0
    free, total = torch.cuda.mem_get_info()
    log(f"[c5 laplace] model set up in {t_setup:.1f} s")
    th = wl.theta_true.copy()
    c, dc, _, _ = wl.params(th[None])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = model.eval(c, dcoefs=dc)
    torch.cuda.synchronize()
    t_eval = time.perf_counter() - t0
    free2, _ = torch.cuda.mem_get_info()
    mlik = float(r["mlik"][0].item())
    grad = r["grad"][0].cpu().numpy()
    var = r["var"][0].cpu().numpy()
    log(f"[c5 laplace] one theta-point with gradient: {t_eval:.1f} s, {int(r['n_iter'][0])} Newton steps, "
        f"mlik {mlik:.6f}")
    res = {"value": 1.0 / t_eval, "unit": "theta-points/s", "s_per_point": t_eval,
           "ccd_27_nodes_s_computed": 27 * t_eval,
           "newton_iters": int(r["n_iter"][0]), "status_ok": bool(r["status"][0] == 0),
           "n": int(g.n), "n_obs": int(n_obs), "nnz_q": int(len(g.indices)), "components": int(g.components.shape[0]),
           "n_theta": int(len(th)), "order": order,
           "mlik": mlik, "grad": [float(v) for v in grad],
           "var_min_max": [float(var.min()), float(var.max())], "setup_s": t_setup,
           "device_free_gb_before_eval": free / 1e9, "device_free_gb_after_eval": free2 / 1e9,
           "device_total_gb": total / 1e9,
           "work": "prior LDLT + selected inversion, Newton to the mode (one LDLT + solve per step), "
                   "posterior selected inversion, mlik, variances, adjoint gradient (one extra solve)",
           "workload": f"spde {nx}x{nx} x rw2(T={T}) + per-slice spde + nugget, n={g.n}, {n_obs} NB(size 5) obs "
                       "from a smooth field, 5 hyperparameters, lean evaluation at theta_true"}
    if fd_components:
        # central differences at two steps and their Richardson extrapolation
        # (g(h) = g + c h^2 + ...): at this size the mlik of an evaluation carries a
        # rounding noise of order 1e-3 (a sum over 1.44M log pivots and 500k terms),
        # which a step of 1e-4 would amplify to O(10)
        fdres = []
        for l in fd_components:
            est = []
            for h in (1e-2, 5e-3):
                vals = []
                for sgn in (h, -h):
                    t = th.copy()
                    t[l] += sgn
                    cc, _, _, _ = wl.params(t[None])
                    t0 = time.perf_counter()
                    rr = model.eval(cc, want_x=False, want_var=False)
                    vals.append(float(rr["mlik"][0].item()))
                    log(f"[c5 laplace] FD evaluation theta[{l}] {sgn:+.0e}: {time.perf_counter() - t0:.1f} s")
                est.append((vals[0] - vals[1]) / (2 * h))
            rich = (4.0 * est[1] - est[0]) / 3.0
            fdres.append({"component": int(l), "fd_h1e-2": est[0], "fd_h5e-3": est[1], "richardson": rich,
                          "adjoint": float(grad[l]), "rel_err_vs_richardson": abs(grad[l] - rich) / max(abs(rich), 1.0)})
        res["gradient_vs_fd"] = fdres
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fit", action="store_true", help="skip the plain-path and end-to-end fit lines")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the Poisson and heterogeneous-batch lines")
    ap.add_argument("--c4", action="store_true",
                    help="add the full-size Besag x AR1 solver line (about 5 GB of HBM)")
    ap.add_argument("--c4-fit", action="store_true",
                    help="add the full-size Besag x AR1 batched evaluation and fit comparison")
    ap.add_argument("--c5", action="store_true",
                    help="add the full-size SPDE x RW2 solver line (about 220 GB of HBM)")
    ap.add_argument("--c5-laplace", action="store_true",
                    help="add the full-size SPDE x RW2 negative binomial Laplace evaluation (lean; minutes)")
    ap.add_argument("--big-order", default="auto", choices=["auto", "chooser", "nd", "amd"],
                    help="ordering of the full-size C4 / C5 lines")
    ap.add_argument("--no-autotune", action="store_true", help="time the default launch schedule")
    ap.add_argument("--stagger", action="store_true",
                    help="let the tuner also try two staggered pipelines of half the batch each (eager, "
                         "half the matrices per launch)")
    ap.add_argument("--cpu-fit", action="store_true",
                    help="add the reference's fit sequence measured on the oracle (host cores; minutes)")
    ap.add_argument("--order", default="auto", choices=["auto", "chooser", "nd", "amd"],
                    help="ordering of the headline and Poisson models (auto: the chooser's ordering or "
                         "nested dissection, whichever has the shorter estimated level schedule)")
    ap.add_argument("--hetero-order", default="chooser", choices=["auto", "chooser", "nd", "amd"],
                    help="ordering of each model of the heterogeneous Laplace lines")
    ap.add_argument("--nx", type=int, default=142)
    ap.add_argument("--n-obs", type=int, default=50_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(theta-points would shard across GPUs with one all-gather of the node outputs).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.stagger:
        os.environ["GMRF_TUNE_STAGGER"] = "1"
    # the tuned schedule goes to a fresh file, so that the line can report it
    if not args.no_autotune and "GMRF_TUNE_FILE" not in os.environ:
        import tempfile

        fd, tune_path = tempfile.mkstemp(prefix="gmrf_schedule_", suffix=".txt")
        os.close(fd)
        os.environ["GMRF_TUNE_FILE"] = tune_path
    from inla_gpu import _lib, ccd
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel
    from inla_gpu.solver import Symbolic

    torch.cuda.init()
    log("== setup: workload")
    wl = W.spde_gaussian_workload(nx=args.nx, n_obs=args.n_obs)
    g = wl.gmrf
    m = 3
    design_z, design_w = ccd.ccd_design(m, 1.0)
    npts = len(design_z)
    stencil = ccd.hessian_points(wl.theta_true, max(1e-4, 0.05 * float(np.max(np.abs(wl.theta_true)))))
    cap = max(npts, len(stencil))
    log(f"== setup: model n={g.n} nnz={len(g.indices)} obs={len(wl.y)} capacity={cap}")
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A,
                         capacity=cap, order=args.order)
    # the prior's pattern equals the posterior's here (A^T A lies inside the SPDE stencil)
    st = Symbolic(g.n, g.indptr, g.indices, order=ORDERS[args.order]).stats()
    # Hessian of the log posterior from one batched evaluation of the stencil
    hh = max(1e-4, 0.05 * float(np.max(np.abs(wl.theta_true))))
    c, dc, lp, dl = wl.params(stencil)
    r = model.eval(c, lpar=lp, want_x=False, want_var=False)
    lpost = r["mlik"].cpu().numpy() + np.array([wl.log_prior(t)[0] for t in stencil])
    H = ccd.hessian_from_values(lpost, m, hh)
    lam, V = ccd.scaling_from_hessian(H)
    thetas = np.array([ccd.theta_from_z(wl.theta_true, lam, V, z) for z in design_z])
    coefs, dcoefs, lpar, dlog = wl.params(thetas)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    coefs, dcoefs, lpar, dlog = dev(coefs), dev(dcoefs), dev(lpar), dev(dlog)

    # the outputs live in fixed buffers, so that every step replays the captured graph
    outs = {"mlik": torch.empty(npts, dtype=torch.float64, device="cuda"),
            "x": torch.empty((npts, g.n), dtype=torch.float64, device="cuda"),
            "var": torch.empty((npts, g.n), dtype=torch.float64, device="cuda"),
            "grad": torch.empty((npts, m), dtype=torch.float64, device="cuda")}

    def step():
        return model.eval(coefs, lpar=lpar, dcoefs=dcoefs, dlog_lpar=dlog, want_x=True, want_var=True, out=outs)

    log(f"== warmup {args.warmup}")
    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    if args.warmup > 0:
        assert (out["status"] == 0).all(), out["status"]
    tune = None
    if not args.no_autotune:
        # the launch schedule tuned on this batch (factorisation + selected inversion of
        # the 2 x 15 prior and posterior matrices), before timing
        step()
        t_def, t_tuned = model.autotune()
        tune = {"default_ms": t_def, "tuned_ms": t_tuned,
                "work": "factorisation + selected inversion of the 30 prior and posterior matrices"}
        try:
            with open(os.environ["GMRF_TUNE_FILE"]) as fh:
                lines = [ln.split() for ln in fh if ln.startswith("gmrf-schedule")]
            if lines:
                f = lines[-1]
                tune["schedule"] = {"batch": int(f[5]), "fuse_below": int(f[6]), "split_max": int(f[7]),
                                    "xcd_max": int(f[8]), "merge_factorisation": f[9], "merge_selinv": f[10],
                                    "stagger": int(f[11]) if len(f) > 11 else -1,
                                    "pp_split_max": int(f[12]) if len(f) > 12 else None,
                                    "trsm32_max": int(f[13]) if len(f) > 13 else None,
                                    "gemm32_max": int(f[14]) if len(f) > 14 else None}
            else:
                tune["schedule"] = "defaults kept (no candidate was faster)"
        except OSError:
            pass
        log(f"== autotune: {t_def:.3f} ms -> {t_tuned:.3f} ms")
        out = step()
        torch.cuda.synchronize()
    log(f"== timed {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = npts * args.steps / (t1 - t0)
    log(f"   {ms_per_step:.3f} ms/step, {value:.1f} theta-points/s")

    # the same timed loop again with per-launch HIP events for the roofline
    log("== profiled pass")
    _lib.profile_enable(True)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    prof = _lib.profile_summary(PEAK_FP64, PEAK_HBM)
    _lib.profile_enable(False)
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    name, d = dom
    avg_s = d["ms"] * 1e-3 / d["launches"]
    fl = d["flops"] / d["launches"]
    by = d["bytes"] / d["launches"]
    # the bound is the roof the kernel's arithmetic intensity (algorithmic flops per
    # algorithmic byte) puts it under: FP64 MFMA above the ridge, HBM below it
    ridge = PEAK_FP64 / PEAK_HBM
    intensity = fl / by if by > 0 else float("inf")
    frac_flop = fl / avg_s / PEAK_FP64
    frac_hbm = by / avg_s / PEAK_HBM
    # the committed counter traffic is per launch of the single pipeline (all 2 x 15
    # matrices); a staggered schedule's launches cover half of them
    staggered = isinstance(tune, dict) and isinstance(tune.get("schedule"), dict) and \
        tune["schedule"].get("stagger", -1) >= 0
    traffic = None if staggered else traffic_per_launch(name)
    if intensity >= ridge:
        roof = {"bound": "mfma", "achieved": fl / avg_s / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": frac_flop, "traffic": traffic}
    else:
        roof = {"bound": "hbm", "achieved": by / avg_s / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": frac_hbm, "traffic": traffic}
    roof.update({"kernel": name, "avg_launch_us": avg_s * 1e6, "flops_per_launch": fl,
                 "algorithmic_bytes_per_launch": by, "intensity_flop_per_byte": intensity,
                 "ridge_flop_per_byte": ridge, "frac_flop": frac_flop, "frac_hbm": frac_hbm})
    if traffic is not None:
        roof["traffic_from"] = os.path.relpath(TRAFFIC_TABLE, ROOT)
        roof["traffic_over_algorithmic"] = traffic / by if by > 0 else None
    # the whole step: the level kernels' time at the roofline, sum over launches of
    # max(F / P_fp64, B / BW), against the measured time per step
    roof_ms = sum(v["roof_ms"] for v in prof.values()) / args.steps
    roof["step_roofline_ms"] = roof_ms
    roof["step_frac"] = roof_ms / ms_per_step
    roof["step_roofline_note"] = "level kernels (factorisation, solves, selected inversion) at max(F/P, B/BW) " \
                                 "per launch, summed, over the measured ms per step"
    kernels = {k: {"launches": v["launches"], "ms_per_step": v["ms"] / args.steps,
                   "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 else 0.0,
                   "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else 0.0,
                   "roof_ms_per_step": v["roof_ms"] / args.steps}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    log(json.dumps(kernels, indent=1))

    grad = out["grad"].cpu().numpy()
    mlik = out["mlik"].cpu().numpy()
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "theta-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"spde-gaussian lattice {args.nx}x{args.nx} n={g.n}, "
                               f"{len(wl.y)} obs, CCD m=3 ({npts} theta-points)",
                   "n": g.n, "n_obs": len(wl.y), "theta_points": npts, "nnz_q": int(len(g.indices)),
                   "ordering": f"{args.order} ({ORDER_KINDS.get(int(st['order_kind']), st['order_kind'])})",
                   "nnz_l": int(st["nnz_l"]), "nsuper": int(st["nsuper"]), "levels": int(st["nlevels"]),
                   "flops_ldlt": st["flops_ldlt"], "flops_selinv": st["flops_selinv"],
                   "mlik_mode": float(mlik[0]), "grad_mode_norm": float(np.linalg.norm(grad[0]))},
        "roofline": roof,
    }
    if tune is not None:
        res["autotune"] = tune
    if not args.no_fit:
        log("== plain per-point GPU path")
        res["plain_gpu"] = plain_gpu(model, wl, thetas)
        res["plain_gpu"]["batched_speedup"] = value / res["plain_gpu"]["value"]
        log("== end-to-end fit: batched BFGS + adjoint vs one-point Nelder-Mead")
        res["fit"], _ = fit_compare(model, wl)
        log(json.dumps(res["fit"]))
    HETERO_ORDER[0] = args.hetero_order
    if not args.no_extra:
        log("== batched Poisson SPDE (CCD m=2)")
        res["poisson_spde"] = poisson_line(args.nx, args.n_obs, order=args.order)
        log(json.dumps(res["poisson_spde"]))
        log("== heterogeneous batch: 64 BYM2 models")
        res["hetero"] = hetero_compare()
        log(json.dumps(res["hetero"]))
        log("== heterogeneous Laplace: 64 BYM2 + Poisson models x 9 CCD points")
        res["hetero_laplace"] = hetero_laplace_line(cpu_sample=not args.no_cpu_baseline)
        log(json.dumps(res["hetero_laplace"]))
        log("== heterogeneous fit: the 64 models' modes, Hessians and CCD nodes as block evaluations")
        res["hetero_fit"] = hetero_fit_line()
        log(json.dumps(res["hetero_fit"]))
    if args.c4:
        log("== full-size Besag x AR1 (n = 200k): factorise, solve, selected inversion")
        res["c4"] = c4_solver_line(order=args.big_order)
        log(json.dumps(res["c4"]))
    if args.c4_fit:
        log("== full-size Besag x AR1 binomial: batched CCD evaluation, BFGS vs Nelder-Mead fit")
        res["c4_fit"] = c4_fit_line(order=args.big_order)
        log(json.dumps(res["c4_fit"]))
    if args.c5:
        log("== full-size SPDE x RW2 (n = 1.44M): factorise, solve, selected inversion")
        res["c5"] = c5_solver_line(order=args.big_order)
        log(json.dumps(res["c5"]))
    if args.c5_laplace:
        log("== full-size SPDE x RW2 negative binomial Laplace, five hyperparameters (n = 1.44M, lean)")
        res["c5_laplace"] = c5_laplace_line(order=args.big_order)
        log(json.dumps(res["c5_laplace"]))
    if args.cpu_fit:
        log("== the reference's fit sequence on the oracle (host cores)")
        res["cpu_fit"] = cpu_fit_line(wl, wl.theta_true + np.array([0.5, -0.3, 0.3]))
        if "fit" in res:
            res["cpu_fit"]["batched_gpu_speedup"] = res["cpu_fit"]["total_s"] / res["fit"]["batched_s"]
        log(json.dumps(res["cpu_fit"]))
    if not args.no_cpu_baseline:
        log(f"== cpu baseline (oracle, {_cpu_cores()} host cores, the {npts} CCD theta-points)")
        res["cpu_baseline"] = cpu_baseline(wl, thetas)
        log(json.dumps(res["cpu_baseline"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
