"""Staggered pipelines of the Gaussian evaluation (the posterior on the model's stream,
the prior on a second stream from a given level of the posterior's factorisation on,
the schedule's `stagger` field; GMRF_STAGGER forces it): every output must equal the
single-pipeline evaluation, which runs the same kernels on the same matrices, for
every stagger level, with and without the gradient and the variances, over repeated
calls; and the oracle's mlik. A version-2 schedule line (with the stagger field) is
read from the tuning file like a version-1 line."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _setup():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload(nx=22, n_obs=900, seed=5)
    g = wl.gmrf
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A, capacity=5)
    return wl, model


def _eval(model, wl, th, grad, var):
    c, dc, lp, dl = wl.params(th)
    kw = {"dcoefs": dc, "dlog_lpar": dl} if grad else {}
    r = model.eval(c, lpar=lp, want_x=True, want_var=var, **kw)
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in r.items() if v is not None}


@pytest.mark.parametrize("grad,var", [(True, True), (False, True), (False, False)])
def test_staggered_evaluation_equals_single_pipeline(monkeypatch, grad, var, record_error):
    wl, model = _setup()
    rng = np.random.default_rng(2)
    th = wl.theta_true[None] + 0.1 * rng.standard_normal((5, 3))
    monkeypatch.delenv("GMRF_STAGGER", raising=False)
    ref = _eval(model, wl, th, grad, var)
    assert (ref["status"] == 0).all()
    keys = [k for k in ("mlik", "x", "var", "grad") if k in ref]
    for level in (0, 1, 3, 6, 1000):  # 1000: beyond the tree, one pipeline
        monkeypatch.setenv("GMRF_STAGGER", str(level))
        for _ in range(2):
            r = _eval(model, wl, th, grad, var)
            assert (r["status"] == 0).all(), (level, r["status"])
            for k in keys:
                assert np.array_equal(r[k], ref[k]), (level, k, np.max(np.abs(r[k] - ref[k])))
    # the oracle at one point (the staggered result equals the single pipeline's)
    c, _, lp, _ = wl.params(th[:1])
    o = orc.find_mode(wl.gmrf.csc(c[0]), "gaussian", "identity", wl.y, lp[0], A=wl.A)
    err = abs(ref["mlik"][0] - o.mlik)
    record_error("mlik (staggered = single)", err, 1e-8, "abs")
    assert err < 1e-8


def test_staggered_evaluation_reports_failed_points(monkeypatch):
    # a point whose prior factor fails (non-finite coefficient) while the others pass
    wl, model = _setup()
    th = np.vstack([wl.theta_true, wl.theta_true + 0.05])
    c, dc, lp, dl = wl.params(th)
    c = c.copy()
    c[1, 0] = np.nan
    monkeypatch.delenv("GMRF_STAGGER", raising=False)
    ref = model.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    monkeypatch.setenv("GMRF_STAGGER", "1")
    r = model.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    assert r["status"][0] == 0 and r["status"][1] != 0
    assert np.array_equal(np.asarray(r["status"]), np.asarray(ref["status"]))
    assert r["mlik"].cpu().numpy()[0] == ref["mlik"].cpu().numpy()[0]
    assert r["mlik"].cpu().numpy()[1] == -np.inf


def test_version2_schedule_line_is_read(tmp_path, monkeypatch):
    import torch

    from inla_gpu import workloads as W
    from inla_gpu.solver import Batch, Symbolic

    g, _, _ = W.spde_components(27, 23)
    c, _ = W.spde_coefs(0.1, 1.2)
    q = sp.csc_matrix(g.csc(c) + 2.0 * sp.identity(g.n, format="csc"))
    q.sort_indices()
    path = tmp_path / "sched.txt"
    monkeypatch.setenv("GMRF_TUNE_FILE", str(path))
    monkeypatch.delenv("GMRF_STAGGER", raising=False)
    sym = Symbolic.from_csc(q)
    st = sym.stats()
    nl = st["nlevels"]
    # count 4: everything merged, no fusion, split-K always, stagger from level 1
    path.write_text(f"gmrf-schedule 2 {st['pattern_hash']:016x} {st['n']} {nl} 4 0 1099511627776 0 "
                    f"{'7' * nl} {'7' * nl} 1\n")
    bt = Batch(sym, 4)
    v = torch.as_tensor(np.tile(q.data, (4, 1)), device="cuda")
    # the tuner, with staggered pipelines of 2 + 2 matrices among its candidates, keeps
    # the results
    monkeypatch.setenv("GMRF_TUNE_STAGGER", "1")
    t_def, t_tuned = bt.autotune(v)
    assert 0 < t_tuned <= t_def
    bt.factorize(v)
    ld, stt = bt.logdet()
    bt.selinv()
    d = bt.selinv_diag().cpu().numpy()
    ref = orc.Factor(q)
    assert (stt == 0).all()
    assert np.max(np.abs(ld - ref.log_abs_det())) <= 1e-10 * abs(ref.log_abs_det())
    assert np.max(np.abs(d - ref.takahashi_diag()[None])) < 1e-9 * np.max(np.abs(d))
