"""Symbolic analysis of the product library (host code, no GPU calls)."""

import numpy as np
import pytest
import scipy.sparse as sp

import matrices as M
from oracle import oracle as orc


def brute_symbolic(q, perm):
    """Elimination tree and column counts by explicit symbolic elimination."""
    n = q.shape[0]
    a = (q.toarray() != 0)[np.ix_(perm, perm)]
    a = a | a.T
    cols = []
    struct = [set(np.nonzero(a[j + 1:, j])[0] + j + 1) for j in range(n)]
    parent = np.full(n, -1)
    for j in range(n):
        s = struct[j]
        if s:
            p = min(s)
            parent[j] = p
            struct[p] |= s - {p}
        cols.append(len(s) + 1)
    return parent, np.array(cols)


CASES = {
    "ar1": M.ar1_q(30, 0.5, 1.0),
    "besag": M.besag_ridge_q(M.FOUR_CYCLE_CHORD, 1.5, 0.25),
    "random": M.random_spd(80, 0.05, 3),
    "random_dense_rows": M.random_spd(70, 0.3, 4),
    "shuffled_lattice": M.shuffled(M.besag_ridge_q([[j for j in (i - 1, i + 1, i - 8, i + 8)
                                                     if 0 <= j < 64 and (abs(j - i) == 8 or j // 8 == i // 8)]
                                                    for i in range(64)], 1.0, 0.1), 1),
    "fgn_like": M.mixture_major(16, 3, 6.0, -0.4, -1.2),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_etree_and_colcounts_match_brute_force(name):
    from inla_gpu.solver import Symbolic

    q = CASES[name]
    sym = Symbolic.from_csc(q)
    perm = sym.perm()
    assert sorted(perm) == list(range(q.shape[0]))
    parent, cc = brute_symbolic(q, perm)
    assert np.array_equal(sym.etree(), parent)
    assert np.array_equal(sym.colcount(), cc)
    st = sym.stats()
    assert st["nnz_l"] == cc.sum()
    # the oracle's simplicial factor with the same permutation has the same fill
    assert orc.Factor(q, perm).nnz_l() == cc.sum()


@pytest.mark.parametrize("name", sorted(CASES))
def test_supernode_partition_is_valid(name):
    from inla_gpu.solver import Symbolic

    q = CASES[name]
    sym = Symbolic.from_csc(q)
    start, spar = sym.supernodes()
    parent = sym.etree()
    n = q.shape[0]
    assert start[0] == 0 and start[-1] == n and np.all(np.diff(start) > 0)
    col2s = np.repeat(np.arange(len(start) - 1), np.diff(start))
    for s in range(len(start) - 1):
        a, b = start[s], start[s + 1]
        for j in range(a, b - 1):
            assert parent[j] == j + 1  # columns of a supernode form a chain
        p = parent[b - 1]
        assert spar[s] == (col2s[p] if p >= 0 else -1)
        assert spar[s] == -1 or spar[s] > s  # postorder


@pytest.mark.parametrize("name", sorted(CASES))
def test_chooser_bit_exact_against_oracle(name):
    from inla_gpu.solver import choose_order

    q = CASES[name]
    k, fwd = choose_order(q)
    ko, fo = orc.choose_order(q)
    assert k == ko
    if k:
        assert np.array_equal(fwd, fo)


def test_chooser_branches():
    from inla_gpu import workloads as W
    from inla_gpu.solver import Symbolic, choose_order

    assert choose_order(M.mixture_major(64, 5, 4.0, -0.3, -1.0))[0] == 1  # time-major
    lat = CASES["shuffled_lattice"]
    assert choose_order(lat)[0] in (0, 2)
    g, _, _ = W.spde_components(40, 40)
    assert choose_order(g.csc([1, 1, 1]))[0] == 0
    # a long shuffled path: RCM halves the envelope
    n = 200
    path = sp.diags([np.ones(n - 1), np.full(n, 3.0), np.ones(n - 1)], [-1, 0, 1])
    ps = M.shuffled(sp.csc_matrix(path), 7)
    k, fwd = choose_order(ps)
    assert k == 2
    assert orc.bandwidth_with_fwd(ps, fwd) == 1
    sym = Symbolic.from_csc(ps)
    assert sym.stats()["order_kind"] == 2
    # the factor order follows the chooser: position w holds the original index inv[w]
    assert np.array_equal(sym.perm(), np.argsort(fwd))


def test_amd_fill_close_to_minimum_degree():
    from inla_gpu import workloads as W
    from inla_gpu.solver import Symbolic

    g, _, _ = W.spde_components(40, 40)
    q = g.csc([1, 1, 1])
    sym = Symbolic(g.n, g.indptr, g.indices)
    md = orc.min_degree(q)
    nnz_md = orc.Factor(q + sp.identity(g.n) * 10, md).nnz_l()
    assert sym.stats()["nnz_l"] < 1.15 * nnz_md


def test_given_and_natural_orders():
    from inla_gpu import _lib
    from inla_gpu.solver import Symbolic

    q = CASES["random"]
    p = np.random.default_rng(0).permutation(q.shape[0])
    sym = Symbolic.from_csc(q, order=_lib.ORDER_GIVEN, perm=p)
    parent, cc = brute_symbolic(q, sym.perm())
    assert np.array_equal(sym.colcount(), cc)
    nat = Symbolic.from_csc(q, order=_lib.ORDER_NATURAL)
    # natural order is postordered, so it is a permutation with the natural fill
    _, ccn = brute_symbolic(q, np.arange(q.shape[0]))
    assert nat.stats()["nnz_l"] == ccn.sum()


def test_invalid_patterns_rejected():
    from inla_gpu.solver import InvalidArgument, Symbolic

    with pytest.raises(InvalidArgument):
        Symbolic(3, [0, 1, 2, 3], [0, 5, 2])
    with pytest.raises(InvalidArgument):
        Symbolic(2, [0, 2, 3], [1, 0, 1])  # unsorted rows
    with pytest.raises(InvalidArgument):
        Symbolic(0, [0], [])
