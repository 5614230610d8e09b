"""The Gaussian evaluation replayed from its captured HIP graph against the same
evaluation run eagerly (the library runs eagerly while per-launch profiling is on).
Replays with changed inputs in the same buffers, with the program's own eager GPU
work and synchronisations between them, and a capture for a second set of buffers:
every replay must equal the eager result (same kernels in the same order)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup():
    from inla_gpu import workloads as W
    from inla_gpu.laplace import LaplaceModel

    wl = W.spde_gaussian_workload(nx=20, n_obs=700, seed=3)
    g = wl.gmrf
    model = LaplaceModel(g.n, g.indptr, g.indices, g.components, wl.y, family="gaussian", A=wl.A, capacity=4)
    return wl, model


def _eager(model, c, lp, dc, dl):
    from inla_gpu import _lib

    _lib.profile_enable(True)
    try:
        r = model.eval(c, lpar=lp, dcoefs=dc, dlog_lpar=dl)
    finally:
        _lib.profile_enable(False)
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in r.items()}


def test_graph_replays_match_eager_evaluation():
    import torch

    wl, model = _setup()
    rng = np.random.default_rng(0)
    B = 3
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    c0, dc0, lp0, dl0 = wl.params(wl.theta_true[None] + 0.1 * rng.standard_normal((B, 3)))
    bufs = [dev(c0), dev(lp0), dev(dc0), dev(dl0)]
    outs = {"mlik": torch.empty(B, dtype=torch.float64, device="cuda"),
            "x": torch.empty((B, wl.n), dtype=torch.float64, device="cuda"),
            "var": torch.empty((B, wl.n), dtype=torch.float64, device="cuda"),
            "grad": torch.empty((B, 3), dtype=torch.float64, device="cuda")}
    for rep in range(5):
        th = wl.theta_true[None] + 0.1 * rng.standard_normal((B, 3))
        c, dc, lp, dl = wl.params(th)
        for t, v in zip(bufs, (c, lp, dc, dl)):
            t.copy_(torch.as_tensor(v, dtype=torch.float64))
        ref = _eager(model, c, lp, dc, dl)
        junk = torch.randn(1 << 20, device="cuda").sum()  # the program's own eager work
        r = model.eval(bufs[0], lpar=bufs[1], dcoefs=bufs[2], dlog_lpar=bufs[3], out=outs)
        torch.cuda.synchronize()
        del junk
        assert (r["status"] == 0).all() and (ref["status"] == 0).all()
        for k in ("mlik", "x", "var", "grad"):
            got = r[k].cpu().numpy()
            assert np.array_equal(got, ref[k]), (rep, k, np.max(np.abs(got - ref[k])))
    # another set of output buffers: a second graph; then the first one again
    r2 = model.eval(bufs[0], lpar=bufs[1], dcoefs=bufs[2], dlog_lpar=bufs[3])
    r1 = model.eval(bufs[0], lpar=bufs[1], dcoefs=bufs[2], dlog_lpar=bufs[3], out=outs)
    for k in ("mlik", "x", "var", "grad"):
        assert np.array_equal(r2[k].cpu().numpy(), r1[k].cpu().numpy())


def test_graph_reports_failed_points():
    # a point whose factors fail (non-finite pivots: Singular) must come back with its
    # status and mlik = -inf from the replayed graph as from the eager path
    import torch

    wl, model = _setup()
    c, dc, lp, dl = wl.params(np.vstack([wl.theta_true, wl.theta_true]))
    c = c.copy()
    c[1, 0] = np.nan
    ref = _eager(model, c, lp, dc, dl)
    cd = torch.as_tensor(c, device="cuda")
    for _ in range(3):
        r = model.eval(cd, lpar=lp, dcoefs=dc, dlog_lpar=dl)
        assert list(r["status"]) == list(ref["status"]) and r["status"][0] == 0 and r["status"][1] != 0
        m = r["mlik"].cpu().numpy()
        assert m[1] == -np.inf and m[0] == ref["mlik"][0]
