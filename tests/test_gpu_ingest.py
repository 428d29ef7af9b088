"""CSR construction on the GPU (gooseberry_amd/csrc/hip/gb_ingest.hip)
against the oracle's reader (the C restatement of src/io.c:369-498, using
the C library's strtoull and strtod): the same nrows, ncols and arrays,
bit for bit, for the same text.  The texts cover what the reader's loop
makes of every kind of line: comments, blank and CRLF lines, a missing
final newline, odd token counts, malformed tokens, signs, hexadecimal and
special values, long digit strings (settled by the exact fallback, more of
them than its first list holds), embedded NUL bytes, 0-based and huge
columns."""
import os
import random
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

from test_oracle_spgemm import random_csr

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def same(got, ref):
    nr, nc, rp, ri, rv = got
    onr, onc, orp, ori, orv = ref
    assert (nr, nc) == (onr, onc)
    assert np.array_equal(rp, orp)
    assert np.array_equal(ri, ori)
    assert np.array_equal(bits(rv), bits(orv))


def check_text(gb, oracle, tmp_path, text, name="t.csr"):
    path = os.path.join(str(tmp_path), name)
    with open(path, "wb") as f:
        f.write(text)
    ref = oracle.read_sparse(path)
    same(gb.read_sparse_device(path), ref)
    same(gb.parse_csr_device(text), ref)
    return ref


def dbl(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


EDGE_LINES = [
    b"", b"1 1.5", b"# comment 3 4", b"% 1 2", b"' 1 2", b"\" 1 2", b"/ 1 2",
    b" # not a comment: a blank-led line", b"\r", b"1 2\r", b"2 3 4", b"5 6.5 7",
    b"3 1.5x 4 2", b"3.7 2.0 9 1", b".5 2", b"-3 2.5", b"+4 -1e-3", b"0 7",
    b"4294967295 1", b"4294967296 1", b"99999999999999999999 2",
    b"1 0x1.8p3 2 0X.8P-2 3 0x 4 0x1p-1074 5 -0x1.fffffffffffffp1023",
    b"1 inf 2 -INFINITY 3 nan 4 -nan(123) 5 nan(0x10) 6 nan(1.5) 7 infinit",
    b"1 1e400 2 -1e-400 3 2.4703282292062328e-324 4 1.7976931348623159e308",
    b"\t1\t2.5\v2\f3.5 ", b"1 2 3", b"7", b"1 2 x 3 4", b"1 1e+ 2 3",
    b"1 2\x003 4", b"\x001 2", b"12 -0 13 -0.0e5",
    b"1 1.00000000000000011102230246251565404236316680908203125",
    b"2 0.30000000000000004 3 123456789012345678901234567890e-10",
]


def test_edge_lines(gb, oracle, tmp_path):
    text = b"\n".join(EDGE_LINES)
    check_text(gb, oracle, tmp_path, text)                     # no final newline
    check_text(gb, oracle, tmp_path, text + b"\n", "nl.csr")   # with one
    check_text(gb, oracle, tmp_path, text.replace(b"\n", b"\r\n") + b"\r\n", "crlf.csr")


@pytest.mark.parametrize("text", [b"", b"\n", b"\n\n\n", b"# only\n% comments",
                                  b"1 2", b"0 0", b"0 1 4294967295 2"],
                         ids=["empty", "nl", "blank3", "comments", "one", "zero", "wrap"])
def test_tiny_texts(gb, oracle, tmp_path, text):
    check_text(gb, oracle, tmp_path, text)


def test_written_matrices(gb, oracle, tmp_path):
    """Files as the reference writes them (%g) and with round-trip digits."""
    a = random_csr(20000, 15000, 8, 31, empty_every=17, dups=True)
    path = os.path.join(str(tmp_path), "g.csr")
    assert oracle.write_csr(path, *a) == 1
    same(gb.read_sparse_device(path), oracle.read_sparse(path))
    rng = random.Random(32)
    lines = []
    rp, ri, rv = a
    for i in range(rp.size - 1):
        toks = []
        for k in range(rp[i], rp[i + 1]):
            f = rng.choice(["%.17g", "%.16e", "%.20e", "%r", "%.3f"])
            v = float(rv[k]) * 10.0 ** rng.randint(-30, 30)
            toks.append("%d %s" % (ri[k] + 1, repr(v) if f == "%r" else f % v))
        lines.append(" ".join(toks))
    check_text(gb, oracle, tmp_path, ("\n".join(lines) + "\n").encode(), "digits.csr")


def test_long_lines_and_many_hard_values(gb, oracle, tmp_path):
    getcontext().prec = 2000
    rng = random.Random(33)
    # halfway points between neighbouring doubles: every one needs the
    # exact fallback; 6000 of them overflow its first list (4096)
    toks = []
    for j in range(6000):
        b = rng.getrandbits(62) | (1 << 61)
        h = (Decimal(dbl(b)) + Decimal(dbl(b + 1))) / 2
        toks.append("%d %s" % (j % 977 + 1, format(h, "e")))
    lines = [" ".join(toks[i:i + 500]) for i in range(0, 6000, 500)]
    # one very long line and many short ones
    lines.append(" ".join("%d %.17g" % (rng.randint(1, 50000), rng.uniform(-1, 1))
                          for _ in range(20000)))
    lines += ["%d %g" % (rng.randint(1, 50000), rng.uniform(-9, 9)) for _ in range(30000)]
    ref = check_text(gb, oracle, tmp_path, "\n".join(lines).encode(), "hard.csr")
    assert ref[2][-1] == 56000


def test_random_bytes(gb, oracle, tmp_path):
    """Text over a small alphabet of the characters the conversions react
    to: whatever the reference's loop makes of it, the GPU makes too."""
    rng = random.Random(34)
    alphabet = b"0123456789 .eE+-xXpPinfaINF()\t\r\n#%\x00_"
    weights = [8] * 10 + [12, 3, 2, 1, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 4,
                          1, 1, 1, 1]
    for t in range(4):
        text = bytes(rng.choices(alphabet, weights=weights, k=200000))
        check_text(gb, oracle, tmp_path, text, "rand%d.csr" % t)


def test_binary_csr_and_resident_read(gb, oracle, tmp_path):
    a = random_csr(5000, 4000, 7, 35, empty_every=9, dups=True)
    csr = os.path.join(str(tmp_path), "m.csr")
    assert oracle.write_csr(csr, *a) == 1
    nr, nc, rp, ri, rv = oracle.read_sparse(csr)
    bcsr = os.path.join(str(tmp_path), "m.bcsr")
    gb.write_csr(bcsr, rp, ri, rv, nc, fmt=gb.FORMAT_BCSR)
    same(gb.read_sparse_device(bcsr, gb.FORMAT_BCSR), (nr, nc, rp, ri, rv))
    same((nr, nc) + gb.read_sparse(bcsr, gb.FORMAT_BCSR)[2:], (nr, nc, rp, ri, rv))
    x = np.random.default_rng(36).uniform(-1, 1, nc)
    ref = oracle.spmult(nr, nc, 1, rp, ri, rv, x)
    for path, fmt in ((csr, gb.FORMAT_CSR), (bcsr, gb.FORMAT_BCSR)):
        M = gb.Matrix.read(path, fmt)
        assert (M.nrows, M.ncols, M.nnz) == (nr, nc, rp[-1])
        y = M.spmv(x)
        assert np.max(np.abs(y - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))
    # a truncated binary file and a wrapped column range are refused
    with open(bcsr, "rb") as f:
        data = f.read()
    bad = os.path.join(str(tmp_path), "bad.bcsr")
    with open(bad, "wb") as f:
        f.write(data[:-8])
    with pytest.raises(gb.GooseberryError) as e:
        gb.read_sparse_device(bad, gb.FORMAT_BCSR)
    assert e.value.name == "INVALIDINPUT"
    wrap = os.path.join(str(tmp_path), "wrap.csr")
    with open(wrap, "wb") as f:
        f.write(b"0 1 4294967295 2\n")
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix.read(wrap)
    assert e.value.name == "INVALIDINPUT"
    with pytest.raises(gb.GooseberryError) as e:
        gb.read_sparse_device(os.path.join(str(tmp_path), "missing.csr"))
    assert e.value.name == "FILENOTFOUND"


def check_mtx(gb, oracle, tmp_path, text, name="m.mtx"):
    path = os.path.join(str(tmp_path), name)
    with open(path, "wb") as f:
        f.write(text)
    ref = oracle.read_sparse(path, mtx=True)
    if isinstance(ref, int):          # the reader refuses the file
        with pytest.raises(gb.GooseberryError) as e:
            gb.read_sparse_device(path, gb.FORMAT_MATRIXMARKET)
        assert e.value.name == "INVALIDINPUT"
        return None
    same(gb.read_sparse_device(path, gb.FORMAT_MATRIXMARKET), ref)
    return ref


def test_matrixmarket_files(gb, oracle, tmp_path):
    """MatrixMarket bodies parsed on the GPU (src/io.c:921-1190): general
    and symmetric files (the mirrored entry first, diagonals twice), files
    without a banner, blank and comment lines, CRLF, long and hard
    weights; and the files the reader refuses."""
    import scipy.io
    import scipy.sparse as sp
    rng = np.random.default_rng(40)
    A = sp.random(700, 500, density=0.01, random_state=41, format="coo")
    A.data = rng.uniform(-1, 1, A.nnz) * 10.0 ** rng.integers(-5, 5, A.nnz)
    p = os.path.join(str(tmp_path), "gen.mtx")
    scipy.io.mmwrite(p, A, precision=17)
    same(gb.read_sparse_device(p, gb.FORMAT_MATRIXMARKET), oracle.read_sparse(p, mtx=True))
    S = sp.random(600, 600, density=0.01, random_state=42)
    S = (S + S.T + sp.eye(600)).tocoo()
    p = os.path.join(str(tmp_path), "sym.mtx")
    scipy.io.mmwrite(p, S, symmetry="symmetric")
    ref = oracle.read_sparse(p, mtx=True)
    same(gb.read_sparse_device(p, gb.FORMAT_MATRIXMARKET), ref)
    assert ref[2][-1] > S.nnz              # mirrored entries and doubled diagonals
    B = b"%%MatrixMarket matrix coordinate real general\n"
    cases = [
        B + b"% c\n\n3 4 5\n1 1 1.5\n\n% mid\n3 4 -2\n2 2 1e-300\n1 4 0x1p3\n3 1 inf\n",
        B.replace(b"\n", b"\r\n") + b"3 3 2\r\n1 2 0.5\r\n3 3 7\r\n",
        b"3 3 3\n1 1\n2 3\n3 2\n",                           # no banner, no weights
        b"%%MatrixMarket matrix coordinate real symmetric\n4 4 3\n2 1 1\n3 3 2\n4 2 3",
        B + b"2 2 1\n1 1 " + b"1" * 30 + b"e-20\n2 1 2.4703282292062327208828439643411068e-324\n",
        B + b"2 2 2\n1 1 1\n2 2\n",                         # weights on some lines only
        B + b"2 2 2\n1 1\n2 2 1\n",
        B + b"2 2 1\n3 1 1\n",                              # row out of range
        B + b"2 2 1\n1 x 1\n",                              # column does not convert
        b"%%MatrixMarket matrix coordinate real symmetric\n3 2 1\n1 3 1\n",
        b"%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
        b"%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
        b"",
        b"\n3 3 1\n1 1 1\n",
        B + b"0 0 0\n",
        B + b"5 5 0\n",
    ]
    refused = 0
    for t, text in enumerate(cases):
        if check_mtx(gb, oracle, tmp_path, text, "c%d.mtx" % t) is None:
            refused += 1
    assert refused >= 6
    # a resident matrix from a MatrixMarket file
    M = gb.Matrix.read(os.path.join(str(tmp_path), "sym.mtx"), gb.FORMAT_MATRIXMARKET)
    assert (M.nrows, M.ncols, M.nnz) == (600, 600, int(ref[2][-1]))
