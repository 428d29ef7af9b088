"""Pins the oracle (oracle/oracle.c) against checks that do not restate it:
brute force on tiny inputs, an independent library (scipy), closed forms of
the Laplacian and textbook CG properties.  CPU only.

The reference ships no golden vectors or known-answer tests for this path
(its test/ directory covers domlib types only), and the reference itself
cannot be built here (its domlib submodule is empty), so these checks are
what pins the oracle.
"""
import math
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def lap2d(nx, ny):
    """Independent scipy construction of the 5-point Dirichlet Laplacian in
    lexicographic order (i = y*nx + x)."""
    Tx = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(nx, nx))
    Ty = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(ny, ny))
    A = sp.kron(sp.identity(ny), Tx) + sp.kron(Ty, sp.identity(nx))
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def lap3d(n):
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(n, n))
    I = sp.identity(n)
    A = sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def rand_csr(n, m, density, seed, longrows=()):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=seed, format="lil")
    for r, L in longrows:
        cols = rng.choice(m, size=min(L, m), replace=False)
        A[r, cols] = 1.0
    A = sp.csr_matrix(A)
    A.data = rng.uniform(-1, 1, A.nnz)
    return A


# ------------------------------------------------------------------ SpMV

def brute_spmv(rowptr, rowind, rowval, x):
    """Row sums left to right from +0.0, each product and sum rounded once
    (Python floats are IEEE doubles; no fused operations)."""
    y = []
    for i in range(len(rowptr) - 1):
        s = 0.0
        for k in range(rowptr[i], rowptr[i + 1]):
            s = s + float(rowval[k]) * float(x[rowind[k]])
        y.append(s)
    return np.array(y)


def test_spmv_bruteforce_bitwise(oracle):
    A = rand_csr(150, 120, 0.05, 3, longrows=[(7, 100), (8, 1)])
    x = np.random.default_rng(1).uniform(-1, 1, 120)
    y = oracle.spmult(150, 120, 1, A.indptr, A.indices, A.data, x)
    assert np.array_equal(y, brute_spmv(A.indptr, A.indices, A.data, x))


def test_spmv_matches_scipy_bitwise(oracle):
    # scipy's csr_matvec also sums each row left to right from 0
    A = rand_csr(3000, 2500, 0.004, 5, longrows=[(0, 2000), (100, 700)])
    x = np.random.default_rng(2).uniform(-1, 1, 2500)
    y = oracle.spmult(3000, 2500, 1, A.indptr, A.indices, A.data, x)
    assert np.array_equal(y, A @ x)


def test_spmv_thread_count_independent(oracle):
    A = lap2d(60, 70)
    x = np.random.default_rng(3).uniform(-1, 1, A.shape[0])
    ys = []
    for t in (1, 3, 8):
        oracle.set_threads(t)
        ys.append(oracle.spmv(A.indptr, A.indices, A.data, x))
    oracle.set_threads(1)
    assert all(np.array_equal(ys[0], y) for y in ys[1:])


def test_spmv_laplacian_times_ones_closed_form(oracle):
    # 2D: row sum of the 5-point stencil = number of out-of-grid neighbours
    nx, ny = 13, 9
    A = lap2d(nx, ny)
    y = oracle.spmv(A.indptr, A.indices, A.data, np.ones(nx * ny))
    xs, ys = np.meshgrid(np.arange(nx), np.arange(ny))
    missing = (xs == 0).astype(int) + (xs == nx - 1) + (ys == 0) + (ys == ny - 1)
    assert np.array_equal(y, missing.ravel().astype(float))


def test_spmv_laplacian_eigenvector(oracle):
    # sin modes are eigenvectors: lambda = sum_d (2 - 2 cos(pi k_d / (N+1)))
    N = 17
    A = lap3d(N)
    k = (2, 5, 3)
    g = np.arange(1, N + 1)
    vx, vy, vz = (np.sin(np.pi * kk * g / (N + 1)) for kk in k)
    v = np.einsum("k,j,i->kji", vz, vy, vx).ravel()  # index (z*N + y)*N + x
    lam = sum(2 - 2 * math.cos(math.pi * kk / (N + 1)) for kk in k)
    y = oracle.spmv(A.indptr, A.indices, A.data, v)
    assert np.max(np.abs(y - lam * v)) < 1e-12 * np.max(np.abs(v)) * 8


def test_spmv_unit_vector_picks_column(oracle):
    A = rand_csr(80, 90, 0.1, 7)
    for j in (0, 45, 89):
        e = np.zeros(90)
        e[j] = 1.0
        y = oracle.spmv(A.indptr, A.indices, A.data, e)
        assert np.array_equal(y, A[:, j].toarray().ravel())


def test_spmm_column_major_b(oracle):
    # src/blas.c:87: bcolval[j*ancols + l]; crowval row-major
    A = rand_csr(40, 30, 0.2, 8)
    B = np.random.default_rng(4).uniform(-1, 1, (30, 3))
    c = oracle.spmult(40, 30, 3, A.indptr, A.indices, A.data, B.T.ravel())
    cols = [brute_spmv(A.indptr, A.indices, A.data, B[:, j]) for j in range(3)]
    assert np.array_equal(c.reshape(40, 3), np.stack(cols, axis=1))


def test_spmult_blocks_cover_only_their_rows(oracle):
    A = rand_csr(50, 50, 0.1, 9)
    x = np.random.default_rng(5).uniform(-1, 1, 50)
    full = oracle.spmv(A.indptr, A.indices, A.data, x)
    c = np.full(50, 123.0)
    oracle.spmult(50, 50, 1, A.indptr, A.indices, A.data, x,
                  blocks=np.array([10, 20, 35]), crowval=c)
    assert np.array_equal(c[10:35], full[10:35])
    assert np.all(c[:10] == 123.0) and np.all(c[35:] == 123.0)


# ------------------------------------------------------------------- dot

def test_dot_sequential_at_one_thread(oracle):
    rng = np.random.default_rng(6)
    a, b = rng.uniform(-1, 1, 1001), rng.uniform(-1, 1, 1001)
    s = 0.0
    for u, v in zip(a, b):
        s = s + float(u) * float(v)
    assert oracle.dot(a, b) == s
    assert abs(oracle.dot(a, b) - math.fsum(a * b)) < 1e-13


# -------------------------------------------------------------------- CG

def test_cg_diagonal_k_distinct_eigenvalues(oracle):
    # CG on an SPD matrix with k distinct eigenvalues converges in k steps
    n = 60
    d = np.repeat([1.0, 3.0, 7.0, 20.0], 15)
    A = sp.csr_matrix(sp.diags(d))
    b = np.random.default_rng(7).uniform(-1, 1, n)
    x, it, rmse, tr = oracle.cgd(A.indptr, A.indices, A.data, b,
                                 error=1e-12 * np.linalg.norm(b), niter=50, trace=True)
    assert it <= 5
    assert np.max(np.abs(x - b / d)) < 1e-12


def test_cg_one_by_one(oracle):
    A = sp.csr_matrix(np.array([[4.0]]))
    x, it, rmse, _ = oracle.cgd(A.indptr, A.indices, A.data, np.array([2.0]), 0.0, niter=3)
    assert x[0] == 0.5 and rmse == 0.0 and it == 1


def test_cg_against_scipy_iterations(oracle):
    # independent solver with a different operation order: same relative
    # residual reached, iteration counts within +-2
    A = lap2d(40, 40)
    b = np.random.default_rng(8).uniform(-1, 1, A.shape[0])
    tol = 1e-8
    x, it, rmse, _ = oracle.cgd(A.indptr, A.indices, A.data, b,
                                error=tol * np.linalg.norm(b), niter=2000)
    count = [0]

    def cb(_):
        count[0] += 1

    xs, info = spla.cg(A, b, rtol=tol, atol=0.0, maxiter=2000, callback=cb)
    assert info == 0
    assert abs(it - count[0]) <= 2
    true_rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true_rel < 2 * tol
    assert rmse <= tol * np.linalg.norm(b)


def test_cg_iteration_cap_and_trace(oracle):
    A = lap2d(30, 30)
    b = np.random.default_rng(9).uniform(-1, 1, A.shape[0])
    x, it, rmse, tr = oracle.cgd(A.indptr, A.indices, A.data, b, 0.0, niter=17, trace=True)
    assert it == 17 and tr.shape == (17, 3)
    # the recurrence residual tracks the true residual
    assert abs(np.linalg.norm(b - A @ x) - rmse) < 1e-9 * np.linalg.norm(b)
    # the residual norm of CG on SPD A is non-increasing in the A^-1 norm,
    # and alpha > 0, beta > 0
    assert np.all(tr[:, 0] > 0) and np.all(tr[:, 1] > 0)


def test_cg_p_update_order_matters_not_for_iterations(oracle):
    # a mistake in alpha/beta (sign, dropped term) stops convergence; the
    # oracle must converge on a random SPD system in <= n iterations
    rng = np.random.default_rng(10)
    M = rng.uniform(-1, 1, (30, 30))
    Ad = M @ M.T + 30 * np.eye(30)
    A = sp.csr_matrix(Ad)
    b = rng.uniform(-1, 1, 30)
    x, it, rmse, _ = oracle.cgd(A.indptr, A.indices, A.data, b, 1e-12, niter=60)
    assert it <= 30
    assert np.allclose(x, np.linalg.solve(Ad, b), atol=1e-10)


# ------------------------------------------------------------- CSR reader

def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


def test_read_csr_semantics(oracle, tmp_path):
    text = "# comment\n1 1.5 3 2.0\n\n2 -1 2 4 1 0.5\n% another\n5 7\n"
    p = _write(tmp_path, "a.csr", text)
    nr, nc, rp, ri, rv = oracle.read_sparse(p)
    assert (nr, nc) == (4, 5)
    assert rp.tolist() == [0, 2, 2, 5, 6]
    assert ri.tolist() == [0, 2, 1, 1, 0, 4]      # file order, duplicates kept
    assert rv.tolist() == [1.5, 2.0, -1.0, 4.0, 0.5, 7.0]


def test_read_csr_zero_based_and_odd_tokens(oracle, tmp_path):
    p = _write(tmp_path, "b.csr", "0 1.0 2 2.0 9\n1 3.0\n")
    nr, nc, rp, ri, rv = oracle.read_sparse(p)
    # a 0 index makes the file 0-based; the dangling "9" is dropped
    assert (nr, nc) == (2, 3)
    assert rp.tolist() == [0, 2, 3] and ri.tolist() == [0, 2, 1]
    assert rv.tolist() == [1.0, 2.0, 3.0]


def test_write_read_csr_roundtrip(oracle, tmp_path):
    A = rand_csr(30, 25, 0.2, 11)
    A.data = np.round(A.data * 8) / 8  # exactly representable in %g
    p = os.path.join(str(tmp_path), "w.csr")
    assert oracle.write_csr(p, A.indptr, A.indices, A.data) == 1
    nr, nc, rp, ri, rv = oracle.read_sparse(p)
    assert nr == 30 and np.array_equal(rp, A.indptr)
    assert np.array_equal(ri, A.indices) and np.array_equal(rv, A.data)


# ------------------------------------------------------- MatrixMarket reader

def test_read_mtx_general_vs_scipy(oracle, tmp_path):
    A = rand_csr(20, 15, 0.2, 12)
    p = os.path.join(str(tmp_path), "g.mtx")
    scipy.io.mmwrite(p, sp.coo_matrix(A), field="real", symmetry="general", precision=17)
    nr, nc, rp, ri, rv = oracle.read_sparse(p, mtx=True)
    B = sp.csr_matrix((rv, ri, rp), shape=(nr, nc))
    assert (B != A).nnz == 0


def test_read_mtx_symmetric_mirror_first_and_diagonal_twice(oracle, tmp_path):
    text = ("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 3\n"
            "1 1 4.0\n2 1 -1.0\n3 3 2.0\n")
    p = _write(tmp_path, "s.mtx", text)
    nr, nc, rp, ri, rv = oracle.read_sparse(p, mtx=True)
    assert rp.tolist() == [0, 3, 4, 6]
    # row 0: mirror of (0,0) then (0,0), then mirror of (1,0) -> (0,1)
    assert ri.tolist() == [0, 0, 1, 0, 2, 2]
    assert rv.tolist() == [4.0, 4.0, -1.0, -1.0, 2.0, 2.0]


def test_read_mtx_no_header_pattern(oracle, tmp_path):
    p = _write(tmp_path, "h.mtx", "2 3 2\n1 3\n2 1\n")
    nr, nc, rp, ri, rv = oracle.read_sparse(p, mtx=True)
    assert (nr, nc) == (2, 3) and ri.tolist() == [2, 0] and rv.tolist() == [1.0, 1.0]


# --------------------------------------------------------------- grid I/O

def test_grid_roundtrip_and_g_format(oracle, tmp_path):
    v = np.array([1.0, -0.5, 1e-20, 123456789.0, 3.14159265358979])
    p = os.path.join(str(tmp_path), "v.vec")
    assert oracle.write_grid(p, v, 5, 1) == 1
    with open(p) as f:
        lines = f.read().splitlines()
    assert lines == ["%g" % t for t in v]
    nr, nc, r = oracle.read_grid(p)
    assert (nr, nc) == (5, 1)
    assert np.array_equal(r, np.array([float("%g" % t) for t in v]))


def test_grid_rejects_ragged(oracle, tmp_path):
    p = _write(tmp_path, "r.grid", "1 2 3\n4 5\n")
    assert oracle.read_grid(p) == 3  # GOOSEBERRY_ERROR_INVALIDINPUT


# ------------------------------------------------------------ index work

@pytest.mark.parametrize("nkeys", [1, 3, 256, 1000])
def test_countingsort_stable_bruteforce(oracle, nkeys):
    rng = np.random.default_rng(nkeys)
    keys = rng.integers(0, nkeys, 3000).astype(np.uint32)
    out, dist = oracle.countingsort_kv(keys, 0, nkeys - 1)
    assert np.array_equal(out, np.argsort(keys, kind="stable"))
    assert np.array_equal(dist, np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=nkeys))]))


def test_permute_sparse_matches_fancy_indexing(oracle):
    A = rand_csr(40, 40, 0.15, 13)
    rng = np.random.default_rng(14)
    rp, cp = rng.permutation(40).astype(np.uint32), rng.permutation(40).astype(np.uint32)
    prp, pri, prv = oracle.permute_sparse(A.indptr, A.indices, A.data, 40, rp, cp)
    P = sp.csr_matrix((prv, pri, prp), shape=(40, 40)).toarray()
    # new row i = old row rp[i]; old column cp[j] becomes column j
    assert np.array_equal(P, A.toarray()[rp][:, cp])


def test_permute_dense_roundtrip(oracle):
    rng = np.random.default_rng(15)
    v = rng.uniform(-1, 1, 12 * 3)
    rp, cp = rng.permutation(12).astype(np.uint32), rng.permutation(3).astype(np.uint32)
    p = oracle.permute_dense(v, 12, 3, rp, cp)
    assert np.array_equal(p.reshape(12, 3), v.reshape(12, 3)[rp][:, cp])
    back = oracle.permute_dense(p, 12, 3, rp, cp, rev=True)
    assert np.array_equal(back, v)


def test_read_labels_and_partition(oracle, tmp_path):
    labels = [2, 0, 1, 2, 0, 0]
    p = _write(tmp_path, "p.part", "# labels\n" + "\n".join(map(str, labels)) + "\n")
    assert oracle.read_labels(p).tolist() == labels
    assert oracle.read_labels(p, nrows=4).tolist() == labels[:4]
    m, perm, dist = oracle.read_partition(p)
    assert m.tolist() == labels
    assert perm.tolist() == np.argsort(labels, kind="stable").tolist()
    assert dist.tolist() == [0, 3, 4, 6]


# ------------------------------------------------------------------ PageRank
# src/gooseberry_bin.c:2622-2737

def _random_graph(n, seed, sink_every=0, deg=6):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if sink_every and i % sink_every == 0:
            rows.append([])
            continue
        cols = rng.choice(n, size=deg, replace=False)
        rows.append(sorted(int(c) for c in cols))
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ri = np.array([c for r in rows for c in r], dtype=np.uint32)
    rv = rng.uniform(0.5, 2.0, ri.size)
    return rp, ri, rv


def test_pagerank_normalisation_is_by_column(oracle):
    rp, ri, rv = _random_graph(400, 1, sink_every=7)
    nval, sinks = oracle.pagerank_prepare(rp, ri, rv, 400)
    A = sp.csr_matrix((nval, ri, rp), shape=(400, 400))
    colsum = np.asarray(A.sum(axis=0)).ravel()
    used = np.bincount(ri, minlength=400) > 0
    assert np.allclose(colsum[used], 1.0, rtol=0, atol=1e-14)
    assert np.array_equal(sinks, np.flatnonzero(np.diff(rp.astype(np.int64)) == 0))


def test_pagerank_without_sinks_solves_the_linear_system(oracle):
    # with no sinks the fixed point is (I - d N) r = (1 - d) / n
    n, d = 300, 0.85
    rp, ri, rv = _random_graph(n, 2)
    nval, sinks = oracle.pagerank_prepare(rp, ri, rv, n)
    assert sinks.size == 0
    out, it, err = oracle.pagerank(rp, ri, nval, sinks, d, 0, 1e-15)
    N = sp.csr_matrix((nval, ri, rp), shape=(n, n))
    r = spla.spsolve((sp.identity(n) - d * N).tocsc(), np.full(n, (1 - d) / n))
    assert np.allclose(out, r, rtol=1e-10, atol=0)
    assert it % 10 == 1 and err <= 1e-15


def test_pagerank_sinks_add_their_whole_mass(oracle):
    # every row empty: out = 0 * d + (1 - d) / n + sum(rank); with two
    # iterations the written vector is the first iterate (the reference's
    # `out` after its last swap)
    n, d = 5, 0.85
    rp = np.zeros(n + 1, dtype=np.uint32)
    ri = np.zeros(0, dtype=np.uint32)
    nval, sinks = oracle.pagerank_prepare(rp, ri, np.zeros(0), n)
    assert list(sinks) == [0, 1, 2, 3, 4]
    out, it, _ = oracle.pagerank(rp, ri, nval, sinks, d, 2, 0.0)
    assert it == 2
    s = 0.0
    for _ in range(n):
        s += 1.0 / n
    assert np.all(out == (1 - d) / n + s)
    assert abs(out[0] - 1.03) < 1e-15
    # one iteration: the written vector is the starting rank
    out1, it1, _ = oracle.pagerank(rp, ri, nval, sinks, d, 1, 0.0)
    assert it1 == 1 and np.all(out1 == 1.0 / n)
