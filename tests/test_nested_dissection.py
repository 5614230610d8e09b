"""Nested dissection and the automatic ordering choice of the product library (host
code, no GPU calls). The ordering has no reference counterpart: the factor it leads to
is checked against brute-force symbolic elimination, and the dissection itself against
properties of the graph (a separator disconnects what it separates, dense rows come
last, components stay contiguous, small graphs keep their order)."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import matrices as M
from test_symbolic import CASES, brute_symbolic


def _sym(q, order):
    from inla_gpu.solver import Symbolic

    return Symbolic.from_csc(q, order=order)


def _lattice(k, diag=4.0):
    from inla_gpu import workloads as W

    adj = W.rook_adjacency(k, k)
    return sp.csc_matrix(W.besag_structure(adj) + diag * sp.identity(k * k))


@pytest.mark.parametrize("name", sorted(CASES))
def test_nd_factor_structure_matches_brute_force(name):
    from inla_gpu import _lib

    q = CASES[name]
    sym = _sym(q, _lib.ORDER_ND)
    perm = sym.perm()
    assert sorted(perm) == list(range(q.shape[0]))
    parent, cc = brute_symbolic(q, perm)
    assert np.array_equal(sym.etree(), parent)
    assert np.array_equal(sym.colcount(), cc)
    assert sym.stats()["nnz_l"] == cc.sum()


def test_small_graph_keeps_its_order():
    from inla_gpu import _lib

    q = CASES["ar1"]  # 30 vertices: at most one leaf of the dissection
    sym = _sym(q, _lib.ORDER_ND)
    assert sym.stats()["order_kind"] == 5
    # the factor order is the postorder of the natural order's tree, i.e. natural
    assert np.array_equal(sym.perm(), np.arange(q.shape[0]))


def test_top_separator_disconnects_the_lattice():
    from inla_gpu import _lib

    k = 40
    q = _lattice(k)
    sym = _sym(q, _lib.ORDER_ND)
    start, spar = sym.supernodes()
    perm = sym.perm()
    roots = np.nonzero(spar == -1)[0]
    assert len(roots) == 1
    # the root supernode (the last columns) holds the top separator: without it the
    # lattice falls apart into at least two pieces, each smaller than half the graph
    # plus the separator
    sep = perm[start[roots[0]]:]
    keep = np.setdiff1d(np.arange(k * k), sep)
    ncomp, lab = connected_components(q[keep][:, keep], directed=False)
    assert ncomp >= 2
    assert np.max(np.bincount(lab)) < (k * k) * 0.75
    assert len(sep) < 3 * k
    # a short, wide tree: far fewer levels than minimum degree's
    amd = _sym(q, _lib.ORDER_AMD)
    assert sym.stats()["nlevels"] < amd.stats()["nlevels"]
    assert sym.stats()["nlevels"] <= 4 * int(np.log2(k * k))


def test_dense_row_is_numbered_last():
    from inla_gpu import _lib

    k = 30
    base = _lattice(k)
    n = k * k + 1
    q = sp.lil_matrix((n, n))
    q[1:, 1:] = base
    q[0, :] = 0.01
    q[:, 0] = 0.01
    q[0, 0] = float(n)
    q = sp.csc_matrix(q)
    q.sort_indices()
    sym = _sym(q, _lib.ORDER_ND)
    assert sym.perm()[-1] == 0
    # the rest is dissected as the lattice alone would be
    assert sym.stats()["nlevels"] <= _sym(base, _lib.ORDER_ND).stats()["nlevels"] + 1


def test_components_stay_contiguous():
    from inla_gpu import _lib

    a, b = _lattice(12), _lattice(9)
    q = sp.csc_matrix(sp.block_diag([a, b]))
    q.sort_indices()
    perm = _sym(q, _lib.ORDER_ND).perm()
    # a component's vertices occupy a contiguous range of factor positions
    for lo, hi in ((0, 144), (144, 144 + 81)):
        pos = np.nonzero((perm >= lo) & (perm < hi))[0]
        assert pos[-1] - pos[0] + 1 == len(pos)


def test_auto_picks_the_shorter_schedule():
    from inla_gpu import _lib
    from inla_gpu import workloads as W

    # SPDE lattice 142 x 142 with 50,000 observations (the headline model's posterior
    # pattern): nested dissection has fewer levels and fewer flops than minimum degree
    wl = W.spde_gaussian_workload(nx=142, n_obs=50_000)
    g = wl.gmrf
    A = sp.csr_matrix(wl.A)
    p = sp.csc_matrix((np.ones(len(g.indices)), g.indices, g.indptr), shape=(g.n, g.n)) + A.T @ A
    p = sp.csc_matrix(p)
    p.sort_indices()
    ch, nd, au = (_sym(p, o) for o in (_lib.ORDER_CHOOSER, _lib.ORDER_ND, _lib.ORDER_AUTO))
    s_ch, s_nd, s_au = ch.stats(), nd.stats(), au.stats()
    assert s_ch["order_kind"] == 0 and s_nd["order_kind"] == 5 and s_au["order_kind"] == 5
    assert np.array_equal(au.perm(), nd.perm())
    assert s_nd["nlevels"] < s_ch["nlevels"]
    assert s_nd["flops_ldlt"] + s_nd["flops_selinv"] < s_ch["flops_ldlt"] + s_ch["flops_selinv"]


@pytest.mark.parametrize("name", sorted(CASES))
def test_auto_is_one_of_the_two(name):
    from inla_gpu import _lib

    q = CASES[name]
    au = _sym(q, _lib.ORDER_AUTO)
    k = au.stats()["order_kind"]
    other = _sym(q, _lib.ORDER_ND if k == 5 else _lib.ORDER_CHOOSER)
    assert np.array_equal(au.perm(), other.perm())
