"""GPU PageRank (src/gooseberry_bin.c:2622-2737) against the oracle.

The reference adds the whole sink mass to every entry
(src/gooseberry_bin.c:2682-2688), so with empty rows present the iteration
grows without bound and the error test never passes: runs with sinks use
a fixed iteration count, runs to an error threshold use sink-free graphs.

The normalisation (column sums in row order, IEEE division), the SpMV of
rows with at most 64 entries, the sink sum (sink order) and the
scale-then-add update are all the reference's operations in the
reference's order, so the written vector is bitwise the oracle's and the
iteration count is equal; only the stopping error is summed in another
order (checked to 1e-12 relative).
"""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_pinning import _random_graph

pytestmark = pytest.mark.gpu


def oracle_pr(oracle, rp, ri, rv, n, d, niter, minerror):
    nval, sinks = oracle.pagerank_prepare(rp, ri, rv, n)
    return oracle.pagerank(rp, ri, nval, sinks, d, niter, minerror) + (sinks.size,)


@pytest.mark.parametrize("sink_every,niter,minerror",
                         [(0, 0, 1e-10), (11, 7, 0.0), (11, 1, 0.0), (0, 40, 1e-6),
                          (11, 25, 0.0)])
def test_pagerank_bitwise_with_oracle(gb, oracle, sink_every, niter, minerror):
    n = 3000
    rp, ri, rv = _random_graph(n, 5, sink_every=sink_every)
    M = gb.Matrix(rp, ri, rv)
    x, it, err, ns = M.pagerank(0.85, niter, minerror)
    xo, ito, erro, nso = oracle_pr(oracle, rp, ri, rv, n, 0.85, niter, minerror)
    assert ns == nso and it == ito
    assert np.array_equal(x, xo)
    assert abs(err - erro) <= 1e-12 * max(erro, 1e-300)


def test_pagerank_long_rows_and_many_sinks(gb, oracle):
    # rows above 64 entries use tree reductions: close, not bitwise
    n = 5000
    rng = np.random.default_rng(8)
    rows = []
    for i in range(n):
        k = 0 if i % 3 == 0 else (300 if i % 97 == 0 else 5)
        rows.append(np.sort(rng.choice(n, size=k, replace=False)))
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ri = np.concatenate(rows).astype(np.uint32)
    rv = rng.uniform(0.1, 1.0, ri.size)
    x, it, err, ns = gb.Matrix(rp, ri, rv).pagerank(0.85, 30, 0.0)
    xo, ito, erro, nso = oracle_pr(oracle, rp, ri, rv, n, 0.85, 30, 0.0)
    assert ns == nso and it == ito == 30
    assert np.max(np.abs(x - xo)) <= 1e-13 * np.max(np.abs(xo))


def test_pagerank_long_columns_bitwise(gb, oracle):
    # two hub columns with thousands of entries (summed by one wavefront
    # each, in row order) among short columns; rows stay short, so the
    # written vector is still bitwise the oracle's
    n = 4000
    rng = np.random.default_rng(12)
    rows = []
    for i in range(n):
        cols = set(rng.choice(n, size=4, replace=False).tolist())
        cols.add(7)
        if i % 2:
            cols.add(1234)
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ri = np.array([c for r in rows for c in r], dtype=np.uint32)
    rv = rng.uniform(0.1, 3.0, ri.size)
    assert np.bincount(ri)[[7, 1234]].min() > 64
    x, it, err, ns = gb.Matrix(rp, ri, rv).pagerank(0.85, 0, 1e-10)
    xo, ito, erro, nso = oracle_pr(oracle, rp, ri, rv, n, 0.85, 0, 1e-10)
    assert ns == nso == 0 and it == ito
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("csplit", [2, 5])
def test_pagerank_column_split_plan(gb, oracle, csplit):
    # the iteration through column slices: each row summed per slice, then
    # the slices added in order -- close to the oracle, with the same
    # iteration count and sinks; the default plan on the same graph stays
    # bitwise
    n = 4000
    rp, ri, rv = _random_graph(n, 7, sink_every=0)
    xo, ito, erro, nso = oracle_pr(oracle, rp, ri, rv, n, 0.85, 0, 1e-9)
    M = gb.Matrix(rp, ri, rv, plan={"csplit": csplit})
    assert M.info()["plan"]["csplit"] == csplit
    x, it, err, ns = M.pagerank(0.85, 0, 1e-9)
    assert ns == nso and abs(it - ito) <= 1
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
    rp2, ri2, rv2 = _random_graph(n, 7, sink_every=9)
    xo2, _, _, nso2 = oracle_pr(oracle, rp2, ri2, rv2, n, 0.85, 12, 0.0)
    x2, it2, _, ns2 = gb.Matrix(rp2, ri2, rv2, plan={"csplit": csplit}).pagerank(0.85, 12, 0.0)
    assert ns2 == nso2 and it2 == 12
    assert np.max(np.abs(x2 - xo2)) <= 1e-12 * np.max(np.abs(xo2))


def test_pagerank_all_sinks(gb):
    n = 5
    rp = np.zeros(n + 1, dtype=np.uint32)
    x, it, _, ns = gb.Matrix(rp, np.zeros(0, np.uint32), np.zeros(0)).pagerank(0.85, 2, 0.0)
    assert ns == 5 and it == 2
    assert np.all(x == (1 - 0.85) / n + 1.0)


def test_pagerank_host_dropin_and_add_scalar(gb, oracle):
    n = 1000
    rp, ri, rv = _random_graph(n, 9)
    x, it, err = gb.pagerank(rp, ri, rv, 0.9, 0, 1e-9)
    xo, ito, erro, _ = oracle_pr(oracle, rp, ri, rv, n, 0.9, 0, 1e-9)
    assert it == ito and np.array_equal(x, xo)
    v = np.linspace(-1, 1, 777)
    assert np.array_equal(gb.add_scalar(v, 0.125), v + 0.125)


def test_pagerank_cli(gb, oracle, tmp_path):
    from gooseberry_amd import CLI_PATH
    n = 2000
    rp, ri, rv = _random_graph(n, 12)
    a = os.path.join(str(tmp_path), "G.csr")
    assert oracle.write_csr(a, rp, ri, rv) == 1
    out = os.path.join(str(tmp_path), "pr.vec")
    r = subprocess.run([CLI_PATH, "pagerank", "-i", a, "-o", out, "-e", "1e-9", "-t"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    # the oracle through the same text input
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    xo, ito, erro, _ = oracle_pr(oracle, orp, ori, orv, nr, 0.85, 0, 1e-9)
    assert "PageRank finished in %d iterations with an RMSE of %f" % (ito, erro) in r.stdout
    assert "Compute:" in r.stdout
    ref = os.path.join(str(tmp_path), "ref.vec")
    oracle.write_grid(ref, xo, nr, 1)
    assert open(out).read() == open(ref).read()
    # with sinks and an iteration cap: the vector before the last iterate
    rp, ri, rv = _random_graph(n, 13, sink_every=9)
    assert oracle.write_csr(a, rp, ri, rv) == 1
    r = subprocess.run([CLI_PATH, "pagerank", "-i", a, "-o", out, "-I", "12", "-d", "0.7"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    xo, ito, erro, _ = oracle_pr(oracle, orp, ori, orv, nr, 0.7, 12, 0.0)
    assert ito == 12 and "PageRank finished in 12 iterations" in r.stdout
    oracle.write_grid(ref, xo, nr, 1)
    assert open(out).read() == open(ref).read()
    # a sparse output file is rejected, as matrix_sparsify does
    r = subprocess.run([CLI_PATH, "pagerank", "-i", a, "-I", "3", "-o",
                        os.path.join(str(tmp_path), "pr.csr")],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 2


def test_pagerank_cli_dense_input(gb, oracle, tmp_path):
    # a dense (grid) matrix: the weights the reference's dense normalisation
    # counts are its nonzero entries (src/gooseberry_bin.c:2633-2644); the
    # build runs the sparse path on them.  (The reference passes the dense
    # matrix's absent row pointers to gooseberry_spmult at :2676, so it has
    # no vector of its own to compare with; the oracle runs the same
    # nonzero entries as CSR.)
    from gooseberry_amd import CLI_PATH
    n = 300
    rng = np.random.default_rng(21)
    W = rng.uniform(0.0, 1.0, (n, n))
    W[rng.uniform(0, 1, (n, n)) < 0.9] = 0.0
    W[5, :] = 0.0  # a sink
    a = os.path.join(str(tmp_path), "W.grid")
    oracle.write_grid(a, W.ravel(), n, n)
    nr, nc, vals = oracle.read_grid(a)
    D = np.asarray(vals).reshape(nr, nc)
    rows, cols = np.nonzero(D)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=nr))]).astype(np.uint32)
    ri, rv = cols.astype(np.uint32), D[rows, cols]
    out = os.path.join(str(tmp_path), "pr.vec")
    r = subprocess.run([CLI_PATH, "pagerank", "-i", a, "-o", out, "-I", "15"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    xo, ito, erro, nsinks = oracle_pr(oracle, rp, ri, rv, nr, 0.85, 15, 0.0)
    assert ito == 15 and nsinks == 1
    ref = os.path.join(str(tmp_path), "ref.vec")
    oracle.write_grid(ref, xo, nr, 1)
    assert open(out).read() == open(ref).read()
    # a non-square dense matrix is refused, as the reference refuses it
    oracle.write_grid(a, W[:, :200].ravel(), n, 200)
    r = subprocess.run([CLI_PATH, "pagerank", "-i", a, "-o", out, "-I", "3"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 2 and "requires a square matrix" in r.stderr


def _hub_graph(n, seed):
    # every row reaches column 7, odd rows column 1234 as well: the chunks
    # the tile walk adds hold many entries of one column
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        cols = set(rng.choice(n, size=4, replace=False).tolist())
        cols.add(7)
        if i % 2:
            cols.add(1234)
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ri = np.array([c for r in rows for c in r], dtype=np.uint32)
    return rp, ri, np.random.default_rng(seed + 1).uniform(0.1, 3.0, ri.size)


def _wide_graph(n, seed):
    # one or two entries a row over more rows than 1024 row blocks of the
    # layout hold: the tile walk takes its groups in two windows
    rng = np.random.default_rng(seed)
    deg = rng.integers(1, 3, n)
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    ri = rng.integers(0, n, int(rp[-1])).astype(np.uint32)
    return rp, ri, rng.uniform(0.5, 2.0, ri.size)


@pytest.mark.parametrize("case", ["hubs", "sinks", "wide"])
def test_pagerank_tile_sums_match_sorted_sums(gb, case, monkeypatch):
    # an x-tile plan sums the columns on its layout, in row order; the sort
    # by column (the path pinned to the oracle above) gives the same sums,
    # so the written vectors are the same bits
    if case == "hubs":
        rp, ri, rv = _hub_graph(20000, 3)
        xtile, niter = 1024, 10
    elif case == "sinks":
        rp, ri, rv = _random_graph(30000, 5, sink_every=7)
        xtile, niter = 4096, 12
    else:
        rp, ri, rv = _wide_graph(8_500_000, 9)
        xtile, niter = 8192, 2
    M = gb.Matrix(rp, ri, rv, plan={"xtile": xtile})
    assert M.info()["plan"]["xtile"] == xtile
    monkeypatch.delenv("GOOSEBERRY_PAGERANK_SORTED_SUMS", raising=False)
    x, it, err, ns = M.pagerank(0.85, niter, 0.0)
    monkeypatch.setenv("GOOSEBERRY_PAGERANK_SORTED_SUMS", "1")
    xs, its, errs, nss = M.pagerank(0.85, niter, 0.0)
    assert (it, ns) == (its, nss) == (niter, ns)
    assert np.array_equal(x, xs)
    assert err == errs
