"""Launch schedules: the autotuner's choice and forced schedules (every small-front
size class merged into one launch per level, no diagonal fusion, split-K always,
no XCD grouping) must give the same factors, solves and selected inverses as the
default schedule (parity bars of the solver tests), and a schedule in the tuning
file is picked up by pattern hash and batch size."""

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _run(sym, vals, count):
    import torch

    from inla_gpu.solver import Batch

    bt = Batch(sym, count)
    v = torch.as_tensor(np.tile(vals, (count, 1)), device="cuda")
    bt.factorize(v)
    ld, st = bt.logdet()
    rhs = torch.as_tensor(np.random.default_rng(1).standard_normal((count, sym.n)), device="cuda")
    x = bt.solve(rhs).cpu().numpy()
    bt.selinv()
    d = bt.selinv_diag().cpu().numpy()
    return ld, st, x, d, rhs.cpu().numpy()


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_forced_schedule_from_tuning_file_matches_default(tmp_path, monkeypatch):
    from inla_gpu import workloads as W
    from inla_gpu.solver import Symbolic

    # a pattern no other test uses (fresh analysis, so the file is read), with small
    # fronts in every size class and large fronts
    g, _, _ = W.spde_components(31, 29)
    c, _ = W.spde_coefs(0.2, 1.4)
    q = sp.csc_matrix(g.csc(c) + 3.0 * sp.identity(g.n, format="csc"))
    q.sort_indices()
    path = tmp_path / "sched.txt"
    monkeypatch.setenv("GMRF_TUNE_FILE", str(path))
    sym = Symbolic.from_csc(q)
    st = sym.stats()
    nl = st["nlevels"]
    # batch size 3: everything merged, no fusion, split-K always, no XCD grouping
    path.write_text(f"gmrf-schedule 1 {st['pattern_hash']:016x} {st['n']} {nl} 3 0 1099511627776 0 "
                    f"{'7' * nl} {'7' * nl}\n")
    ld3, st3, x3, d3, rhs3 = _run(sym, q.data, 3)    # the forced schedule
    ld2, st2, x2, d2, rhs2 = _run(sym, q.data, 2)    # the defaults
    assert (st3 == 0).all() and (st2 == 0).all()
    assert abs(ld3[0] - ld2[0]) <= 1e-12 * abs(ld2[0])
    assert _rel(x3[:2], x2) < 1e-12
    assert _rel(d3[:2], d2) < 1e-12
    ref = orc.Factor(q)
    assert abs(ld3[0] - ref.log_abs_det()) <= 1e-10 * abs(ref.log_abs_det())
    assert _rel(x3[0], ref.solve(rhs3[0])) < 1e-10
    assert _rel(d3[0], ref.takahashi_diag()) < 1e-9


def test_autotune_keeps_results(monkeypatch, tmp_path):
    import torch

    from inla_gpu.solver import Batch, Symbolic

    monkeypatch.setenv("GMRF_TUNE_FILE", str(tmp_path / "tuned.txt"))
    q = M.random_spd(900, 0.012, 11)
    sym = Symbolic.from_csc(q)
    count = 8
    ld0, _, x0, d0, _ = _run(sym, q.data, count)
    bt = Batch(sym, count)
    v = torch.as_tensor(np.tile(q.data, (count, 1)), device="cuda")
    t_def, t_tuned = bt.autotune(v)
    print(f"autotune: default {t_def:.3f} ms, tuned {t_tuned:.3f} ms")
    assert 0 < t_tuned <= t_def
    ld1, st1, x1, d1, _ = _run(sym, q.data, count)
    assert (st1 == 0).all()
    assert np.max(np.abs(ld1 - ld0)) <= 1e-12 * np.max(np.abs(ld0))
    assert _rel(x1, x0) < 1e-12 and _rel(d1, d0) < 1e-12


@pytest.mark.parametrize("pp_split", [0, 1 << 40])
def test_sigma_pp_split_schedule_matches_default(tmp_path, monkeypatch, pp_split):
    # a version-3 schedule line: Sigma[P,P] never split (0) or always split into three
    # 32 x 32 workgroups; factors, solves and selected inverses equal the other form's
    # and the oracle's
    from inla_gpu import workloads as W
    from inla_gpu.solver import Symbolic

    g, _, _ = W.spde_components(33, 26)
    c, _ = W.spde_coefs(-0.2, 1.1)
    q = sp.csc_matrix(g.csc(c) + 1.5 * sp.identity(g.n, format="csc"))
    q.sort_indices()
    path = tmp_path / "sched.txt"
    monkeypatch.setenv("GMRF_TUNE_FILE", str(path))
    sym = Symbolic.from_csc(q)
    st = sym.stats()
    nl = st["nlevels"]
    path.write_text(f"gmrf-schedule 3 {st['pattern_hash']:016x} {st['n']} {nl} 3 2048 96 256 "
                    f"{'0' * nl} {'0' * nl} -1 {pp_split}\n")
    ld3, st3, x3, d3, rhs3 = _run(sym, q.data, 3)    # the line's schedule
    ld2, st2, x2, d2, rhs2 = _run(sym, q.data, 2)    # the defaults (split up to 1024 tasks)
    assert (st3 == 0).all() and (st2 == 0).all()
    assert abs(ld3[0] - ld2[0]) <= 1e-12 * abs(ld2[0])
    assert _rel(x3[:2], x2) < 1e-12
    assert _rel(d3[:2], d2) < 1e-12
    ref = orc.Factor(q)
    assert _rel(d3[0], ref.takahashi_diag()) < 1e-9


@pytest.mark.parametrize("t32", [0, 1 << 40])
def test_tile32_schedules_match_default(tmp_path, monkeypatch, t32):
    # a version-4 schedule line: panel solves and Sigma[B,P] never (0) or always as four
    # 32 x 32 workgroups per tile, Sigma[P,P] likewise never or always split; factors,
    # solves and selected inverses equal the defaults' and the oracle's
    from inla_gpu import workloads as W
    from inla_gpu.solver import Symbolic

    g, _, _ = W.spde_components(29, 34)
    c, _ = W.spde_coefs(0.3, 1.3)
    q = sp.csc_matrix(g.csc(c) + 2.5 * sp.identity(g.n, format="csc"))
    q.sort_indices()
    path = tmp_path / "sched.txt"
    monkeypatch.setenv("GMRF_TUNE_FILE", str(path))
    sym = Symbolic.from_csc(q)
    st = sym.stats()
    nl = st["nlevels"]
    path.write_text(f"gmrf-schedule 4 {st['pattern_hash']:016x} {st['n']} {nl} 3 2048 96 256 "
                    f"{'0' * nl} {'0' * nl} -1 {t32} {t32} {t32}\n")
    ld3, st3, x3, d3, _ = _run(sym, q.data, 3)    # the line's schedule
    ld2, st2, x2, d2, _ = _run(sym, q.data, 2)    # the defaults
    assert (st3 == 0).all() and (st2 == 0).all()
    assert abs(ld3[0] - ld2[0]) <= 1e-12 * abs(ld2[0])
    assert _rel(x3[:2], x2) < 1e-12
    assert _rel(d3[:2], d2) < 1e-12
    ref = orc.Factor(q)
    assert abs(ld3[0] - ref.log_abs_det()) <= 1e-10 * abs(ref.log_abs_det())
    assert _rel(d3[0], ref.takahashi_diag()) < 1e-9
