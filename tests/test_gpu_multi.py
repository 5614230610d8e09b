"""Heterogeneous batches (gmrf_multi_*): k patterns factorised as their
block-diagonal union in one launch sequence. Bars as for single patterns:
log|Q| 1e-10 relative, solves 1e-10 relative, selected-inverse diagonals 1e-9
relative; per-block results also match single-pattern batches."""

import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def _mats():
    from inla_gpu import workloads as W

    rng = np.random.default_rng(11)
    adj = W.delaunay_adjacency(300, rng)
    R = W.besag_structure(adj)
    bym = sp.csc_matrix(0.5 * sp.identity(300) + 0.5 * R + sp.identity(300) * np.e)  # + Poisson weights
    bym.sort_indices()
    path = sp.diags([np.ones(199), np.full(200, 3.0), np.ones(199)], [-1, 0, 1])
    out = []
    for q in [M.random_spd(120, 0.05, 3), M.ar1_q(500, 0.9, 2.0), bym,
              M.mixture_major(64, 5, 4.0, -0.3, -1.0),          # chooser: time-major
              M.shuffled(sp.csc_matrix(path), 7),                 # chooser: RCM
              M.besag_ridge_q(M.FOUR_CYCLE_CHORD, 1.5, 0.25)]:
        q = sp.csc_matrix(q)
        q.sum_duplicates()
        q.sort_indices()
        out.append(q)
    return out


def test_multi_matches_oracle_and_single_batches():
    import torch

    from inla_gpu.solver import Batch, MultiBatch, Symbolic, choose_order

    mats = _mats()
    mb = MultiBatch(mats, capacity=2)
    assert list(mb.order_kind) == [choose_order(q)[0] for q in mats]
    v = np.concatenate([q.data for q in mats])
    vals = torch.as_tensor(np.stack([v, 2.0 * v]), device="cuda")
    mb.factorize(vals)
    ld, st = mb.logdet()
    assert np.all(st == 0)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(mb.n)
    x = mb.solve(torch.as_tensor(np.stack([b, b]), device="cuda")).cpu().numpy()
    mb.selinv()
    dg = mb.selinv_diag().cpu().numpy()
    for q, Q in enumerate(mats):
        a, e = mb.n_off[q], mb.n_off[q + 1]
        f = orc.Factor(Q)
        ref_ld = f.log_abs_det()
        n = Q.shape[0]
        assert abs(ld[0, q] - ref_ld) < 1e-10 * abs(ref_ld) + 1e-12
        assert abs(ld[1, q] - (ref_ld + n * np.log(2.0))) < 1e-10 * abs(ref_ld) + 1e-9
        xs = np.linalg.solve(Q.toarray(), b[a:e])
        assert _rel(x[0, a:e], xs) < 1e-10
        assert _rel(x[1, a:e], xs / 2.0) < 1e-10
        di = np.diag(np.linalg.inv(Q.toarray()))
        assert _rel(dg[0, a:e], di) < 1e-9
        assert _rel(dg[1, a:e], di / 2.0) < 1e-9
        # the same block through its own single-pattern batch
        bt = Batch(Symbolic.from_csc(Q), 1)
        bt.factorize(torch.as_tensor(Q.data[None], device="cuda"))
        lds, _ = bt.logdet()
        assert abs(lds[0] - ld[0, q]) < 1e-11 * max(1.0, abs(lds[0]))


def test_scale_factors_reference_values_on_gpu():
    from inla_gpu import workloads as W

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "scale_model.json")))
    structs, kinds = [], []
    for c in fx["cases"]:
        if c["kind"] == "besag":
            structs.append(W.besag_structure(c["adj"]))
            kinds.append("constant")
        else:
            structs.append(W.rw2_structure(c["n"]))
            kinds.append("linear")
    fac = W.scale_factors(structs, kinds)
    for c, R, f in zip(fx["cases"], structs, fac):
        exact = np.exp(np.mean(np.log(np.diag(np.linalg.pinv(R.toarray())))))
        assert f == pytest.approx(exact, rel=1e-9)
        got = f * R[0, 0] if c["expect"] == "scaled diagonal" else f
        assert got == pytest.approx(c["value"], rel=1e-6)
