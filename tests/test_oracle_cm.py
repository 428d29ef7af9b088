"""Pinning of the oracle's Cuthill-McKee ordering (src/permute.c:49-110,
288-332) on the CPU.

The reference orders equal degrees by domlib's priority queue, whose source
is not in the reference tree; the oracle (and the GPU) take equal degrees in
ascending vertex order.  Where no tie arises the ordering is the textbook
one, so it is checked against scipy's independent implementation
(scipy.sparse.csgraph.reverse_cuthill_mckee, reversed): its start is the
least-degree vertex of each component and it sorts each vertex's newly
found neighbours by degree with a stable insertion sort, which on sorted
rows is the same (degree, vertex) order.  The graphs are built so that the
least-degree vertex of every component is unique.  Closed forms: a path is
left in place, a star starts at its first leaf.  Blocks reorder each
diagonal block on its own.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee


def csr_of(adj):
    a = sp.csr_matrix(adj)
    a.sort_indices()
    return a.indptr.astype(np.uint32), a.indices.astype(np.uint32)


def random_graph(n, extra, seed, comps=1):
    """Symmetric, with a diagonal, connected components of about n/comps
    vertices; every component has one vertex of least degree."""
    rng = np.random.default_rng(seed)
    A = sp.lil_matrix((n, n), dtype=np.float64)
    bounds = np.linspace(0, n, comps + 1).astype(int)
    for c in range(comps):
        lo, hi = bounds[c], bounds[c + 1]
        order = rng.permutation(np.arange(lo, hi))
        for a, b in zip(order[:-1], order[1:]):       # a spanning path
            A[a, b] = A[b, a] = 1
        for _ in range(extra * (hi - lo)):
            a, b = rng.integers(lo, hi, 2)
            if a != b:
                A[a, b] = A[b, a] = 1
        # one vertex of least degree: give every other vertex of the
        # component at least one more neighbour than the chosen one
        A = A.tocsr().tolil()
        deg = np.asarray((A[lo:hi, lo:hi] != 0).sum(axis=1)).ravel()
        star = lo + int(np.argmin(deg))
        dmin = deg.min()
        for v in range(lo, hi):
            if v == star:
                continue
            while (A[v, lo:hi] != 0).sum() <= dmin:
                w = int(rng.integers(lo, hi))
                if w != v and w != star:
                    A[v, w] = A[w, v] = 1
        # the chosen vertex may have gained neighbours: keep it the least
        deg = np.asarray((A[lo:hi, lo:hi] != 0).sum(axis=1)).ravel()
        assert (deg == deg.min()).sum() == 1
    A.setdiag(4.0)
    return A.tocsr()


def test_path_is_left_in_place(oracle):
    n = 60
    A = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(n, n), format="csr")
    rp, ri = csr_of(A)
    assert np.array_equal(oracle.cuthillmckee(rp, ri), np.arange(n))


def test_star_starts_at_its_first_leaf(oracle):
    m = 9
    A = sp.lil_matrix((m + 1, m + 1))
    A.setdiag(1)
    for k in range(1, m + 1):
        A[0, k] = A[k, 0] = 1
    rp, ri = csr_of(A)
    assert list(oracle.cuthillmckee(rp, ri)) == [1, 0] + list(range(2, m + 1))


@pytest.mark.parametrize("n,extra,comps,seed", [(40, 1, 1, 1), (300, 2, 1, 2),
                                                (500, 3, 4, 3), (2000, 2, 7, 4)])
def test_matches_scipy_without_ties_of_starts(oracle, n, extra, comps, seed):
    A = random_graph(n, extra, seed, comps)
    rp, ri = csr_of(A)
    ref = np.asarray(reverse_cuthill_mckee(A, symmetric_mode=True))[::-1]
    got = oracle.cuthillmckee(rp, ri)
    # scipy visits the components in order of its own (unstable) degree
    # argsort; compare component by component, in start order by degree
    assert sorted(got.tolist()) == list(range(n))
    if comps == 1:
        assert np.array_equal(got, ref)
    else:
        lab = np.zeros(n, dtype=int)
        bounds = np.linspace(0, n, comps + 1).astype(int)
        for c in range(comps):
            lab[bounds[c]:bounds[c + 1]] = c
        for c in range(comps):
            assert np.array_equal(got[lab[got] == c], ref[lab[ref] == c])


def test_isolated_and_empty_rows_first_by_degree(oracle):
    # rows 3 and 7 have no entries (degree 0) and come first, in vertex
    # order; row 5 has only its diagonal (degree 1); then the path 0-1-2
    n = 8
    A = sp.lil_matrix((n, n))
    for a, b in ((0, 1), (1, 2), (4, 6)):
        A[a, b] = A[b, a] = 1
    for v in (0, 1, 2, 4, 5, 6):
        A[v, v] = 1
    rp, ri = csr_of(A)
    got = oracle.cuthillmckee(rp, ri)
    assert list(got[:3]) == [3, 7, 5]
    assert sorted(got.tolist()) == list(range(n))
    assert list(got) == [3, 7, 5, 0, 1, 2, 4, 6]


def test_blocks_reorder_each_diagonal_block(oracle):
    A = random_graph(600, 2, 5, comps=1)
    rp, ri = csr_of(A)
    blocks = np.array([0, 150, 151, 400, 600], dtype=np.uint32)
    got = oracle.cuthillmckee(rp, ri, blocks=blocks)
    for lo, hi in zip(blocks[:-1], blocks[1:]):
        sub = A[lo:hi, lo:hi].tocsr()
        srp, sri = csr_of(sub)
        assert np.array_equal(got[lo:hi], lo + oracle.cuthillmckee(srp, sri))


def test_bandwidth_of_a_scrambled_grid_is_restored(oracle):
    nx = 30
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(nx, nx))
    G = (sp.kron(sp.eye(nx), T) + sp.kron(T, sp.eye(nx))).tocsr()
    q = np.random.default_rng(6).permutation(nx * nx)
    S = G[q][:, q].tocsr()
    rp, ri = csr_of(S)
    perm = oracle.cuthillmckee(rp, ri)
    P = S[perm][:, perm].tocoo()
    assert np.max(np.abs(P.row - P.col)) <= nx + 1
    Sc = S.tocoo()
    assert np.max(np.abs(Sc.row - Sc.col)) > 10 * nx
