"""Distinct handles used concurrently from several threads, as the reference runs one
solver per CCD node inside a Rayon par_iter (crates/inla_stats/src/inference.rs:1336-1337,
1394; FaerCpuSolver is Send + Sync with &mut self): every thread owns its handles, some
share a pattern (and so the process-wide analysis), and each result must equal what
the same handle computes alone, and the oracle's. The library's calls release the
interpreter lock (ctypes), so the threads overlap on the host and on the device."""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _cases():
    from inla_gpu import workloads as W

    g, _, _ = W.spde_components(21, 19)
    qs = []
    for lt, lk in ((0.0, 1.2), (0.4, 0.8)):  # one pattern, two value sets
        c, _ = W.spde_coefs(lt, lk)
        q = sp.csc_matrix(g.csc(c) + 0.5 * sp.identity(g.n, format="csc"))
        q.sort_indices()
        qs.append(q)
    qs.append(M.random_spd(700, 0.01, 9))
    return qs


def test_solvers_in_threads_match_single_thread_and_oracle(record_error):
    from inla_gpu.solver import HipSolver

    qs = _cases()
    rng = np.random.default_rng(0)
    rhs = [rng.standard_normal(q.shape[0]) for q in qs]
    refs = []
    for q, b in zip(qs, rhs):
        f = orc.Factor(q)
        refs.append((f.log_abs_det(), f.solve(b), f.takahashi_diag()))
    alone = []
    for q, b in zip(qs, rhs):
        s = HipSolver()
        s.factorize(q)
        alone.append((s.log_abs_det(), s.solve(b), s.diag_inv()))

    def work(t):
        s = HipSolver()
        out = []
        for it in range(6):
            k = (t + it) % len(qs)
            s.factorize(qs[k])
            out.append((k, s.log_abs_det(), s.solve(rhs[k]), s.diag_inv()))
        return out

    with ThreadPoolExecutor(8) as ex:
        results = list(ex.map(work, range(8)))
    err = 0.0
    for out in results:
        for k, ld, x, d in out:
            a_ld, a_x, a_d = alone[k]
            assert ld == a_ld and np.array_equal(x, a_x) and np.array_equal(d, a_d), k
            r_ld, r_x, r_d = refs[k]
            err = max(err, abs(ld - r_ld) / abs(r_ld), float(np.max(np.abs(x - r_x)) / np.max(np.abs(r_x))),
                      float(np.max(np.abs(d - r_d)) / np.max(np.abs(r_d))))
    record_error("8 threads x 6 factorisations vs oracle", err, 1e-9)
    assert err < 1e-9


def test_models_in_threads_match_single_thread():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_poisson_workload(nx=12, n_obs=500, seed=6)
    g = wl.gmrf
    th = wl.theta_true[None] + np.array([[0.0, 0.0], [0.15, -0.1], [-0.1, 0.2]])
    c, dc, _, _ = wl.params(th)

    def make():
        return LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="poisson", A=wl.A, capacity=3)

    def run(model):
        r = model.eval(c, dcoefs=dc)
        return {k: r[k].cpu().numpy() for k in ("mlik", "x", "var", "grad")}, np.asarray(r["status"])

    ref, st = run(make())
    assert (st == 0).all()
    models = [make() for _ in range(4)]

    def work(i):
        return [run(models[i]) for _ in range(3)]

    with ThreadPoolExecutor(4) as ex:
        results = list(ex.map(work, range(4)))
    for out in results:
        for r, s in out:
            assert (s == 0).all()
            for k in ref:
                assert np.array_equal(r[k], ref[k]), k
