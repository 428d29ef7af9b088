"""The other sparse formats of src/io.c: METIS/Chaco graph (:242-367),
CLUTO (:501-611), libsvm (:612-745) and coordinate points (:746-920), and
the graph, MatrixMarket and point writers (:1277-1462), through
gooseberry_read_sparse_matrix / gooseberry_write_sparse_matrix.

Pinned three ways: small files whose arrays are worked out by hand from the
reference's statements; the oracle's Python restatement (which calls the C
library's own conversions) on generated files; round trips through the
writers.  Where the reference would write past its arrays or store an index
wrapped below zero or past the matrix, the library refuses the file with
GOOSEBERRY_ERROR_INVALIDINPUT and the oracle raises."""
import os

import numpy as np
import pytest


def write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "wb") as f:
        f.write(text.encode() if isinstance(text, str) else text)
    return p


def read(gb, path, fmt):
    return gb.read_sparse(path, fmt)


def check(got, nrows, ncols, rowptr, rowind, rowval):
    n, m, rp, ri, rv = got
    assert (n, m) == (nrows, ncols)
    assert rp.tolist() == rowptr
    assert ri.tolist() == rowind
    assert rv.tolist() == rowval


GOLDEN = {
    # header "n m": 2m entries, no weights; 1-based neighbours
    "plain.graph": ("% comment\n3 3\n2 3\n1 3\n1 2\n", 5,
                    (3, 3, [0, 2, 4, 6], [1, 2, 0, 2, 0, 1], [1.0] * 6)),
    # weights (fmt 1), a comment between vertices, an empty last line = a
    # vertex without edges
    "weighted.graph": ("4 2 1\n2 0.5\n1 0.5 3 7\n# not a vertex\n2 7\n\n", 5,
                       (4, 4, [0, 1, 3, 4, 4], [1, 0, 2, 1], [0.5, 0.5, 7.0, 7.0])),
    # header "nrows ncols nnz"; empty lines are skipped
    "m.clu": ("2 4 3\n1 1.5 4 2\n\n3 -1\n", 6,
              (2, 4, [0, 2, 3], [0, 3, 2], [1.5, 2.0, -1.0])),
    # label then index:value; an index ending its line ends the row
    "m.svm": ("+1 1:0.5 3:2\n-1 2:1.5\n# c\n0 4:1 5\n", 4,
              (3, 4, [0, 2, 3, 4], [0, 2, 1, 3], [0.5, 2.0, 1.5, 1.0])),
    # rows 1,2,1 and columns 1,1,2: minrow = mincol = 1; grouped by row in
    # file order
    "m.ij": ("1 1 2.0\n2 1 3.0\n1 2 4.0\n", 3,
             (2, 2, [0, 2, 3], [0, 1, 0], [2.0, 4.0, 3.0])),
}


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_files(gb, oracle, tmp_path, name):
    text, fmt, want = GOLDEN[name]
    p = write(tmp_path, name, text)
    check(read(gb, p, fmt), *want)
    ref = {gb.FORMAT_GRAPH: oracle.read_graph, gb.FORMAT_CLU: oracle.read_clu,
           gb.FORMAT_SVM: oracle.read_svm, gb.FORMAT_POINT: oracle.read_point}[fmt]
    check(ref(p), *want)


REFUSED = [
    ("more vertex lines than the header", "g.graph", "2 1\n2\n1\n1\n"),
    ("more edges than the header", "g.graph", "2 1\n2 2\n1\n"),
    ("vertex 0", "g.graph", "2 1\n0\n1\n"),
    ("vertex past the vertex lines", "g.graph", "2 1\n3\n1\n"),
    ("no header", "g.graph", "% only a comment\n"),
    ("CLUTO column past ncols", "m.clu", "1 2 1\n3 1.0\n"),
    ("CLUTO more rows than the header", "m.clu", "1 2 2\n1 1.0\n2 1.0\n"),
    ("libsvm index 0", "m.svm", "1 0:1.0\n"),
    ("point without a value", "m.ij", "1 2\n"),
    # columns 1, 3, 2: the last non-maximum (2) becomes mincol, so column 1
    # would wrap below zero
    ("point index base", "m.ij", "1 1 1\n1 3 1\n2 2 1\n"),
]


@pytest.mark.parametrize("what,name,text", REFUSED, ids=[r[0] for r in REFUSED])
def test_refused_files(gb, oracle, tmp_path, what, name, text):
    fmt = {".graph": gb.FORMAT_GRAPH, ".clu": gb.FORMAT_CLU, ".svm": gb.FORMAT_SVM,
           ".ij": gb.FORMAT_POINT}[os.path.splitext(name)[1]]
    p = write(tmp_path, name, text)
    with pytest.raises(gb.GooseberryError) as e:
        read(gb, p, fmt)
    assert e.value.name == "INVALIDINPUT"
    ref = {gb.FORMAT_GRAPH: oracle.read_graph, gb.FORMAT_CLU: oracle.read_clu,
           gb.FORMAT_SVM: oracle.read_svm, gb.FORMAT_POINT: oracle.read_point}[fmt]
    with pytest.raises(ValueError):
        ref(p)


def small_matrix():
    # row 0 holds a loop; weights 5, 2, 2, 0.25, 0.25
    return (np.array([0, 2, 4, 5], np.uint32), np.array([0, 1, 0, 2, 1], np.uint32),
            np.array([5.0, 2.0, 2.0, 0.25, 0.25]))


def test_graph_writer_text(gb, oracle, tmp_path):
    rp, ri, rv = small_matrix()
    p = os.path.join(str(tmp_path), "w.graph")
    gb.write_csr(p, rp, ri, rv, 3, gb.FORMAT_GRAPH)
    # one row holds a loop: 4 entries, 2 edges; weights differ: fmt 01;
    # integral weights as integers
    want = "3 2 01\n 2 2\n 1 2 3 0.25\n 2 0.25\n"
    assert open(p).read() == want
    q = os.path.join(str(tmp_path), "o.graph")
    oracle.write_graph(q, rp, ri, rv, 3)
    assert open(q).read() == want
    gb.write_csr(p, rp, ri, np.ones(5), 3, gb.FORMAT_GRAPH)
    assert open(p).read() == "3 2\n 2\n 1 3\n 2\n"
    with pytest.raises(gb.GooseberryError):
        gb.write_csr(p, rp, ri, rv, 4, gb.FORMAT_GRAPH)   # not square


def test_matrixmarket_and_point_writer_text(gb, oracle, tmp_path):
    rp = np.array([0, 1, 2], np.uint32)
    ri = np.array([2, 0], np.uint32)
    rv = np.array([1.5, -2.0])
    p = os.path.join(str(tmp_path), "w.mtx")
    gb.write_csr(p, rp, ri, rv, 3, gb.FORMAT_MATRIXMARKET)
    want = ("%%MatrixMarket matrix coordinate real general\n%\n"
            "% Written by Gooseberry version 0.1.7\n"
            "%   http://cs.umn.edu/~lasalle/gooseberry\n%\n"
            "2 3 2\n1\t3\t1.500000\n2\t1\t-2.000000\n")
    assert open(p).read() == want
    q = os.path.join(str(tmp_path), "o.mtx")
    oracle.write_mtx(q, rp, ri, rv, 3)
    assert open(q).read() == want
    # the MatrixMarket reader reads it back
    n, m, rp2, ri2, rv2 = gb.read_sparse(p, gb.FORMAT_MATRIXMARKET)
    assert (n, m) == (2, 3) and rp2.tolist() == [0, 1, 2] and ri2.tolist() == [2, 0]
    assert rv2.tolist() == [1.5, -2.0]
    p = os.path.join(str(tmp_path), "w.ij")
    gb.write_csr(p, rp, ri, rv, 3, gb.FORMAT_POINT)
    assert open(p).read() == "1\t3\t1.500000\n2\t1\t-2.000000\n"


def test_point_reader_cannot_read_general_point_output(gb, tmp_path):
    # the reference's index-base rule: with two entries in the last row the
    # repeated row number becomes minrow, so its own writer's output for a
    # general matrix is refused (where the reference would wrap row 1)
    rp, ri, rv = small_matrix()
    p = os.path.join(str(tmp_path), "w.ij")
    gb.write_csr(p, rp, ri, rv, 3, gb.FORMAT_POINT)
    with pytest.raises(gb.GooseberryError):
        gb.read_sparse(p, gb.FORMAT_POINT)
    # a diagonal matrix keeps both bases at 1 and round-trips
    d = np.arange(1, 6, dtype=np.float64) * 0.5
    gb.write_csr(p, np.arange(6, dtype=np.uint32), np.arange(5, dtype=np.uint32), d, 5,
                 gb.FORMAT_POINT)
    n, m, rp2, ri2, rv2 = gb.read_sparse(p, gb.FORMAT_POINT)
    assert (n, m) == (5, 5) and ri2.tolist() == list(range(5)) and rv2.tolist() == d.tolist()


def random_symmetric(n, deg, seed, quarters=True):
    rng = np.random.default_rng(seed)
    A = {}
    for i in range(n):
        for j in rng.choice(n, size=deg, replace=False):
            if i != j:
                v = float(rng.integers(1, 4000)) / (4.0 if quarters else 1.0)
                A[(i, int(j))] = v
                A[(int(j), i)] = v
    rows = [[] for _ in range(n)]
    for (i, j), v in sorted(A.items()):
        rows[i].append((j, v))
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return (rp, np.array([j for r in rows for j, _ in r], np.uint32),
            np.array([v for r in rows for _, v in r]))


def test_graph_round_trip(gb, oracle, tmp_path):
    rp, ri, rv = random_symmetric(300, 4, 1)
    p = os.path.join(str(tmp_path), "r.graph")
    gb.write_csr(p, rp, ri, rv, 300, gb.FORMAT_GRAPH)
    n, m, rp2, ri2, rv2 = gb.read_sparse(p, gb.FORMAT_GRAPH)
    assert (n, m) == (300, 300)
    assert np.array_equal(rp2, rp) and np.array_equal(ri2, ri) and np.array_equal(rv2, rv)
    o = oracle.read_graph(p)
    assert np.array_equal(o[2], rp) and np.array_equal(o[3], ri) and np.array_equal(o[4], rv)


def generated_files(seed):
    """Files of each format with comments, blank lines, CR line ends,
    signs, exponents and odd spacing."""
    rng = np.random.default_rng(seed)
    num = lambda: rng.choice(["%d" % rng.integers(-9, 10), "%.3e" % rng.normal(),
                              "%g" % rng.uniform(-5, 5), "+%d.5" % rng.integers(0, 9)])
    sp = lambda: rng.choice([" ", "  ", "\t", " \t"])
    eol = lambda: rng.choice(["\n", "\r\n"])
    n = 40
    out = {}
    # graph, weighted
    lines = ["%d %d 1" % (n, n * 3)]
    for i in range(n):
        if rng.random() < 0.1:
            lines.append("# comment")
        k = rng.integers(0, 5)
        lines.append(sp().join("%d%s%s" % (rng.integers(1, n + 1), sp(), num())
                               for _ in range(k)))
    out[("g.graph", 5)] = eol().join(lines) + "\n"
    # clu
    lines = ["%d %d %d" % (n, 50, n * 6)]
    for i in range(n):
        if rng.random() < 0.1:
            lines.append("")
        k = rng.integers(1, 6)
        lines.append(sp().join("%d %s" % (rng.integers(1, 51), num()) for _ in range(k)))
    out[("m.clu", 6)] = eol().join(lines) + "\n"
    # svm
    lines = []
    for i in range(n):
        if rng.random() < 0.1:
            lines.append("% c")
        k = rng.integers(0, 6)
        lines.append(("%+d" % rng.integers(-1, 2)) + "".join(
            "%s%d:%s" % (sp(), rng.integers(1, 70), num()) for _ in range(k)))
    out[("m.svm", 4)] = eol().join(lines) + "\n"
    # points, rows ascending with one entry each, columns ascending: bases 1
    lines = ["%d%s%d%s%s" % (i + 1, sp(), 2 * i + 1, sp(), num()) for i in range(n)]
    out[("m.ij", 3)] = eol().join(lines) + "\n"
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_generated_files_match_oracle(gb, oracle, tmp_path, seed):
    for (name, fmt), text in generated_files(seed).items():
        p = write(tmp_path, name, text)
        ref = {5: oracle.read_graph, 6: oracle.read_clu, 4: oracle.read_svm,
               3: oracle.read_point}[fmt]
        try:
            want = ref(p)
        except ValueError:
            with pytest.raises(gb.GooseberryError):
                gb.read_sparse(p, fmt)
            continue
        got = gb.read_sparse(p, fmt)
        assert got[:2] == want[:2], name
        for a, b in zip(got[2:], want[2:]):
            assert np.array_equal(a.view(np.uint8), np.asarray(b).view(np.uint8)), name
