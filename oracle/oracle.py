"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (oracle.c), the
plain C restatement of the reference's sparse-solve path.  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker or the timed CPU baseline -- never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_LIB = None
_P = C.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(LIB_PATH)
        u32, f64, sz, i32 = C.c_uint32, C.c_double, C.c_size_t, C.c_int
        L.orc_spmult.argtypes = [u32, u32, u32, _P, _P, _P, _P, _P, _P, u32]
        L.orc_spmult.restype = i32
        L.orc_dot.argtypes = [u32, _P, _P]
        L.orc_dot.restype = f64
        L.orc_cgd.argtypes = [u32, u32, _P, _P, _P, u32, _P, u32, _P, f64, sz, _P, _P, sz,
                              C.POINTER(sz), C.POINTER(f64)]
        L.orc_cgd.restype = i32
        for nm in ("orc_read_csr", "orc_read_mtx"):
            getattr(L, nm).argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32),
                                       C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]
            getattr(L, nm).restype = i32
        L.orc_read_grid.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32), C.POINTER(_P)]
        L.orc_read_grid.restype = i32
        L.orc_write_grid.argtypes = [C.c_char_p, u32, u32, _P]
        L.orc_write_grid.restype = i32
        L.orc_write_csr.argtypes = [C.c_char_p, u32, _P, _P, _P]
        L.orc_write_csr.restype = i32
        L.orc_read_labels.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(_P)]
        L.orc_read_labels.restype = i32
        L.orc_countingsort_kv.argtypes = [_P, _P, u32, u32, sz, _P, _P]
        L.orc_countingsort_kv.restype = None
        L.orc_read_partition.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32),
                                         C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]
        L.orc_read_partition.restype = i32
        L.orc_permute_sparse.argtypes = [u32, u32, _P, _P, _P, _P, _P]
        L.orc_permute_sparse.restype = i32
        L.orc_permute_dense.argtypes = [u32, u32, _P, _P, _P]
        L.orc_permute_dense.restype = i32
        L.orc_permute_dense_rev.argtypes = [u32, u32, _P, _P, _P]
        L.orc_permute_dense_rev.restype = i32
        L.orc_spmultsp.argtypes = [u32, u32, u32, _P, _P, _P, _P, _P, _P, _P]
        L.orc_spmultsp.restype = i32
        L.orc_spmultsp_sp.argtypes = [u32, u32, u32, _P, _P, _P, _P, _P, _P,
                                      C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]
        L.orc_spmultsp_sp.restype = i32
        L.orc_add_scalar.argtypes = [u32, _P, f64]
        L.orc_add_scalar.restype = i32
        L.orc_pagerank_prepare.argtypes = [u32, u32, _P, _P, _P, _P, C.POINTER(u32)]
        L.orc_pagerank_prepare.restype = i32
        L.orc_pagerank.argtypes = [u32, _P, _P, _P, _P, u32, f64, sz, f64, _P,
                                   C.POINTER(sz), C.POINTER(f64)]
        L.orc_pagerank.restype = i32
        L.orc_cuthillmckee.argtypes = [u32, u32, _P, _P, _P, u32, _P]
        L.orc_cuthillmckee.restype = i32
        L.orc_free.argtypes = [_P]
        L.orc_set_num_threads.argtypes = [i32]
        L.orc_num_threads.restype = i32
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def set_threads(t):
    lib().orc_set_num_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


def spmult(anrows, ancols, bncols, rowptr, rowind, rowval, bcolval, blocks=None,
           crowval=None):
    rowptr, rowind, rowval, bcolval = _u32(rowptr), _u32(rowind), _f64(rowval), _f64(bcolval)
    if crowval is None:
        crowval = np.zeros(anrows * bncols)
    blocks = _u32(blocks)
    nb = 0 if blocks is None else blocks.size - 1
    rv = lib().orc_spmult(anrows, ancols, bncols, _p(rowptr), _p(rowind), _p(rowval),
                          _p(bcolval), _p(crowval), _p(blocks), nb)
    assert rv == 1
    return crowval


def spmv(rowptr, rowind, rowval, x):
    n = len(rowptr) - 1
    return spmult(n, len(x), 1, rowptr, rowind, rowval, x)


def dot(a, b):
    a, b = _f64(a), _f64(b)
    return lib().orc_dot(a.size, _p(a), _p(b))


_DEVNULL = None


def _devnull():
    """A C stdio stream on /dev/null (for the reference's per-iteration
    print, src/cgd.c:116, when the solve is timed)."""
    global _DEVNULL
    if _DEVNULL is None:
        c = C.CDLL(None)
        c.fopen.restype = C.c_void_p
        c.fopen.argtypes = [C.c_char_p, C.c_char_p]
        _DEVNULL = c.fopen(b"/dev/null", b"w")
    return _DEVNULL


def cgd(rowptr, rowind, rowval, b, error, niter=0, trace=False, print_devnull=False):
    """Returns (x, iterations, final rmse, trace array (k x 3) or None).
    print_devnull: print each iteration's line to /dev/null, as the
    reference prints it to stdout."""
    rowptr, rowind, rowval, b = _u32(rowptr), _u32(rowind), _f64(rowval), _f64(b)
    n = rowptr.size - 1
    x = np.zeros(n)
    cap = int(niter) if (trace and niter) else (20000 if trace else 0)
    tr = np.zeros(3 * cap) if cap else None
    it = C.c_size_t()
    rm = C.c_double()
    log = _devnull() if print_devnull else None
    rv = lib().orc_cgd(n, n, _p(rowptr), _p(rowind), _p(rowval), n, _p(b), n, _p(x),
                       float(error), int(niter), log, _p(tr), cap, C.byref(it), C.byref(rm))
    assert rv == 1, rv
    k = it.value
    return x, k, rm.value, (tr[:3 * min(k, cap)].reshape(-1, 3) if cap else None)


def _take(p, n, ct, dtype):
    if n == 0:
        lib().orc_free(p)
        return np.zeros(0, dtype=dtype)
    a = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()
    lib().orc_free(p)
    return a


def read_sparse(path, mtx=False):
    nr, nc = C.c_uint32(), C.c_uint32()
    rp, ri, rv = _P(), _P(), _P()
    f = lib().orc_read_mtx if mtx else lib().orc_read_csr
    code = f(path.encode(), C.byref(nr), C.byref(nc), C.byref(rp), C.byref(ri), C.byref(rv))
    if code != 1:
        return code
    rowptr = _take(rp, nr.value + 1, C.c_uint32, np.uint32)
    nnz = int(rowptr[-1])
    return (nr.value, nc.value, rowptr, _take(ri, nnz, C.c_uint32, np.uint32),
            _take(rv, nnz, C.c_double, np.float64))


def read_grid(path):
    nr, nc = C.c_uint32(), C.c_uint32()
    rv = _P()
    code = lib().orc_read_grid(path.encode(), C.byref(nr), C.byref(nc), C.byref(rv))
    if code != 1:
        return code
    return nr.value, nc.value, _take(rv, nr.value * nc.value, C.c_double, np.float64)


def write_grid(path, vals, nrows, ncols):
    vals = _f64(vals)
    return lib().orc_write_grid(path.encode(), nrows, ncols, _p(vals))


def write_csr(path, rowptr, rowind, rowval):
    rowptr, rowind, rowval = _u32(rowptr), _u32(rowind), _f64(rowval)
    return lib().orc_write_csr(path.encode(), rowptr.size - 1, _p(rowptr), _p(rowind),
                               _p(rowval))


def read_labels(path, nrows=0):
    n = C.c_uint32(nrows)
    p = _P()
    code = lib().orc_read_labels(path.encode(), C.byref(n), C.byref(p))
    if code != 1:
        return code
    return _take(p, n.value, C.c_uint32, np.uint32)


def countingsort_kv(keys, kmin, kmax, vals=None):
    keys = _u32(keys)
    vals = np.arange(keys.size, dtype=np.uint32) if vals is None else _u32(vals)
    out = np.empty(keys.size, dtype=np.uint32)
    dist = np.empty(int(kmax) - int(kmin) + 2, dtype=np.uint32)
    lib().orc_countingsort_kv(_p(keys), _p(vals), kmin, kmax, keys.size, _p(out), _p(dist))
    return out, dist


def read_partition(path, nrows=0):
    n, k = C.c_uint32(nrows), C.c_uint32()
    mp, pp, dp = _P(), _P(), _P()
    code = lib().orc_read_partition(path.encode(), C.byref(n), C.byref(k), C.byref(mp),
                                    C.byref(pp), C.byref(dp))
    if code != 1:
        return code
    return (_take(mp, n.value, C.c_uint32, np.uint32), _take(pp, n.value, C.c_uint32, np.uint32),
            _take(dp, k.value + 1, C.c_uint32, np.uint32))


def permute_sparse(rowptr, rowind, rowval, ncols, rowperm=None, colperm=None):
    rowptr, rowind, rowval = _u32(rowptr).copy(), _u32(rowind).copy(), _f64(rowval).copy()
    lib().orc_permute_sparse(rowptr.size - 1, ncols, _p(rowptr), _p(rowind), _p(rowval),
                             _p(_u32(rowperm)), _p(_u32(colperm)))
    return rowptr, rowind, rowval


def permute_dense(vals, nrows, ncols, rowperm=None, colperm=None, rev=False):
    vals = _f64(vals).copy()
    f = lib().orc_permute_dense_rev if rev else lib().orc_permute_dense
    f(nrows, ncols, _p(vals), _p(_u32(rowperm)), _p(_u32(colperm)))
    return vals


def pagerank_prepare(rowptr, rowind, rowval, ncols):
    """Normalised values (column sums in row order) and the sink rows."""
    rowptr, rowind = _u32(rowptr), _u32(rowind)
    nval = np.array(rowval, dtype=np.float64, copy=True)
    n = rowptr.size - 1
    sinks = np.zeros(max(n, 1), dtype=np.uint32)
    ns = C.c_uint32()
    lib().orc_pagerank_prepare(n, ncols, _p(rowptr), _p(rowind), _p(nval), _p(sinks),
                               C.byref(ns))
    return nval, sinks[:ns.value].copy()


def pagerank(rowptr, rowind, nval, sinks, damping=0.85, niter=0, minerror=0.0):
    """Returns (out vector, iterations, rmse) as the reference prints them."""
    rowptr, rowind, nval, sinks = _u32(rowptr), _u32(rowind), _f64(nval), _u32(sinks)
    n = rowptr.size - 1
    out = np.zeros(n, dtype=np.float64)
    it, err = C.c_size_t(), C.c_double()
    lib().orc_pagerank(n, _p(rowptr), _p(rowind), _p(nval), _p(sinks), sinks.size,
                       damping, niter, minerror, _p(out), C.byref(it), C.byref(err))
    return out, it.value, err.value


def spmultsp(a, b, bncols):
    """Dense C = A B for sparse A, B given as (rowptr, rowind, rowval)."""
    (arp, ari, arv), (brp, bri, brv) = a, b
    arp, ari, arv = _u32(arp), _u32(ari), _f64(arv)
    brp, bri, brv = _u32(brp), _u32(bri), _f64(brv)
    n = arp.size - 1
    out = np.zeros(n * bncols, dtype=np.float64)
    lib().orc_spmultsp(n, brp.size - 1, bncols, _p(arp), _p(ari), _p(arv), _p(brp),
                       _p(bri), _p(brv), _p(out))
    return out


def spmultsp_sp(a, b, bncols):
    """Sparse C = A B: (rowptr, rowind, rowval) in the reference's order."""
    (arp, ari, arv), (brp, bri, brv) = a, b
    arp, ari, arv = _u32(arp), _u32(ari), _f64(arv)
    brp, bri, brv = _u32(brp), _u32(bri), _f64(brv)
    n = arp.size - 1
    cp, ci, cv = _P(), _P(), _P()
    lib().orc_spmultsp_sp(n, brp.size - 1, bncols, _p(arp), _p(ari), _p(arv), _p(brp),
                          _p(bri), _p(brv), C.byref(cp), C.byref(ci), C.byref(cv))
    rowptr = _take(cp, n + 1, C.c_uint32, np.uint32)
    nnz = int(rowptr[-1])
    return rowptr, _take(ci, nnz, C.c_uint32, np.uint32), _take(cv, nnz, C.c_double, np.float64)


def cuthillmckee(rowptr, rowind, ncols=None, blocks=None):
    """perm[p] = the row placed at position p (src/permute.c:49-110,288-332)."""
    rowptr, rowind = _u32(rowptr), _u32(rowind)
    n = rowptr.size - 1
    perm = np.full(n, 0xffffffff, dtype=np.uint32)
    b = None if blocks is None else _u32(blocks)
    code = lib().orc_cuthillmckee(n, n if ncols is None else ncols, _p(rowptr), _p(rowind),
                                  _p(b), 0 if b is None else b.size - 1, _p(perm))
    if code != 1:
        raise ValueError("orc_cuthillmckee: %d" % code)
    return perm


# ---------------------------------------------------------------------------
# The other sparse formats of src/io.c, restated in Python for small files.
# Number conversions call the C library's strtoull/strtod/strtoll on the
# line's bytes, so end positions and values are the reference's.  Lines are
# split at '\n' with trailing '\r' removed; where the reference would write
# past its arrays or store a wrapped index, these raise ValueError.

_LIBC = None


def _libc():
    global _LIBC
    if _LIBC is None:
        c = C.CDLL(None)
        for nm, rt in (("strtoull", C.c_ulonglong), ("strtoll", C.c_longlong),
                       ("strtod", C.c_double)):
            f = getattr(c, nm)
            f.restype = rt
            f.argtypes = ([C.c_void_p, C.POINTER(C.c_void_p)] if nm == "strtod"
                          else [C.c_void_p, C.POINTER(C.c_void_p), C.c_int])
        _LIBC = c
    return _LIBC


class _Line:
    """One NUL-terminated line with the C conversions at byte offsets."""

    def __init__(self, text):
        self.buf = C.create_string_buffer(text)
        self.base = C.addressof(self.buf)
        self.n = len(text)

    def _call(self, nm, off, *extra):
        end = C.c_void_p()
        v = getattr(_libc(), nm)(C.c_void_p(self.base + off), C.byref(end), *extra)
        return v, end.value - self.base

    def ull(self, off):
        v, e = self._call("strtoull", off, 10)
        return v & 0xFFFFFFFF, e, v

    def ll(self, off):
        return self._call("strtoll", off, 10)

    def dbl(self, off):
        return self._call("strtod", off)

    def at(self, off):
        return self.buf.raw[off:off + 1] if off <= self.n else b""


def _lines(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n")
    if parts and parts[-1] == b"" and data.endswith(b"\n"):
        parts.pop()
    return [p.rstrip(b"\r\n") for p in parts] if data else []


def _is_comment(line):
    return line[:1] in (b"#", b"%", b"'", b'"', b"/")


def _csr(rows, ncols):
    rowptr = np.zeros(len(rows) + 1, dtype=np.uint32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    ri = np.array([c for r in rows for c, _ in r], dtype=np.uint32)
    rv = np.array([v for r in rows for _, v in r], dtype=np.float64)
    return len(rows), ncols, rowptr, ri, rv


def read_graph(path):
    """src/io.c:242-367: (nrows, ncols, rowptr, rowind, rowval)."""
    lines = _lines(path)
    k = 0
    while k < len(lines) and (lines[k] == b"" or _is_comment(lines[k])):
        k += 1
    hn = cap = wgt = 0
    if k < len(lines):
        L = _Line(lines[k])
        hn, e, _ = L.ull(0)
        m, e2, _ = L.ull(e)
        cap = (m * 2) & 0xFFFFFFFF
        _, e3, wfull = L.ull(e2)
        # (int) of the conversion: its low 32 bits decide
        wgt = (wfull & 0xFFFFFFFF) if e3 != e2 else 0
    if hn == 0 and cap == 0:
        raise ValueError("empty graph")
    rows, nnz = [], 0
    for line in lines[k + 1:]:
        if line != b"" and _is_comment(line):
            continue
        if len(rows) == hn:
            raise ValueError("more vertex lines than the header")
        L = _Line(line)
        row = []
        i, e, _ = L.ull(0)
        s = 0
        f = 0.0
        if wgt:
            s = e
            f, e = L.dbl(s)
        while e != s:
            if nnz == cap:
                raise ValueError("more edges than the header")
            if i == 0:
                raise ValueError("vertex 0")
            row.append((i - 1, f if wgt else 1.0))
            nnz += 1
            s = e
            i, e, _ = L.ull(s)
            if wgt:
                s = e
                f, e = L.dbl(s)
        rows.append(row)
    n = len(rows)
    if any(c >= n for r in rows for c, _ in r):
        raise ValueError("vertex beyond the vertex lines")
    return _csr(rows, n)


def read_clu(path):
    """src/io.c:501-611."""
    lines = [l for l in _lines(path)]
    k = 0
    while k < len(lines) and (lines[k] == b"" or _is_comment(lines[k])):
        k += 1
    if k == len(lines):
        raise ValueError("no header")
    fields = []
    L = _Line(lines[k])
    off = 0
    for _ in range(3):
        # sscanf %u: a conversion that fails stops the rest
        v, e, _ = L.ull(off)
        if e == off:
            break
        fields.append(v)
        off = e
    hn, maxcol, cap = (fields + [0, 0, 0])[:3]
    rows, nnz = [], 0
    for line in lines[k + 1:]:
        if line == b"" or _is_comment(line):
            continue
        if len(rows) == hn:
            raise ValueError("more rows than the header")
        L = _Line(line)
        i, e, _ = L.ull(0)
        s = e
        f, e = L.dbl(s)
        row = []
        while e != s:
            if nnz == cap:
                raise ValueError("more entries than the header")
            if i == 0 or i > maxcol:
                raise ValueError("column outside 1..ncols")
            row.append((i - 1, f))
            nnz += 1
            s = e
            i, e, _ = L.ull(s)
            s = e
            f, e = L.dbl(s)
        rows.append(row)
    return _csr(rows, maxcol)


def read_svm(path):
    """src/io.c:612-745 (a row ends where an index ends its line)."""
    rows, maxcol = [], 0
    for line in _lines(path):
        if line == b"" or _is_comment(line):
            continue
        L = _Line(line)
        _, e = L.ll(0)
        s = e
        row = []
        while True:
            i, e, _ = L.ull(s)
            if L.at(e) in (b"\x00", b""):
                break
            s = e + 1
            f, e = L.dbl(s)
            if e == s:
                break
            if i == 0:
                raise ValueError("index 0")
            maxcol = max(maxcol, i)
            row.append((i - 1, f))
            s = e
        rows.append(row)
    return _csr(rows, maxcol)


def read_point(path):
    """src/io.c:746-920, with the reference's index-base rule."""
    ents = []
    nrows = ncols = 0
    minrow = mincol = 1
    for line in _lines(path):
        if line == b"" or _is_comment(line):
            continue
        L = _Line(line)
        i, e, _ = L.ull(0)
        j, e, _ = L.ull(e)
        s = e
        w, e = L.dbl(s)
        if e == s:
            raise ValueError("error in point list at line %d" % len(ents))
        if i > nrows:
            nrows = i
        else:
            minrow = i
        if j > ncols:
            ncols = j
        else:
            mincol = j
        ents.append((i, j, w))
    nrows = (nrows + 1 - minrow) & 0xFFFFFFFF
    ncols = (ncols + 1 - mincol) & 0xFFFFFFFF
    rows = [[] for _ in range(nrows)]
    for i, j, w in ents:
        r, c = (i - minrow) & 0xFFFFFFFF, (j - mincol) & 0xFFFFFFFF
        if r >= nrows or c >= ncols:
            raise ValueError("entry outside the matrix")
        rows[r].append((c, w))
    return _csr(rows, ncols)


def write_graph(path, rowptr, rowind, rowval, ncols):
    """src/io.c:1277-1366 (text)."""
    n = len(rowptr) - 1
    if n != ncols:
        raise ValueError("not square")
    nnz = int(rowptr[n])
    for i in range(n):
        if any(int(rowind[j]) == i for j in range(rowptr[i], rowptr[i + 1])):
            nnz -= 1
    ew = rowval is not None and any(rowval[j] != rowval[j - 1] for j in range(1, nnz))
    out = []
    if ew:
        out.append("%d %d 01\n" % (ncols, nnz // 2))
        for i in range(n):
            s = ""
            for j in range(rowptr[i], rowptr[i + 1]):
                c, v = int(rowind[j]), float(rowval[j])
                if c == i:
                    continue
                if -2.0 ** 63 <= v < 2.0 ** 63 and v == float(int(v)):
                    s += " %d %d" % (c + 1, int(v))
                else:
                    s += " %d %g" % (c + 1, v)
            out.append(s + "\n")
    else:
        out.append("%d %d\n" % (ncols, nnz // 2))
        for i in range(n):
            out.append("".join(" %d" % (int(rowind[j]) + 1)
                               for j in range(rowptr[i], rowptr[i + 1])
                               if int(rowind[j]) != i) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_mtx(path, rowptr, rowind, rowval, ncols):
    """src/io.c:1367-1418 (text)."""
    n = len(rowptr) - 1
    out = ["%%MatrixMarket matrix coordinate real general\n", "%\n",
           "% Written by Gooseberry version 0.1.7\n",
           "%   http://cs.umn.edu/~lasalle/gooseberry\n", "%\n",
           "%d %d %d\n" % (n, ncols, int(rowptr[n]))]
    for i in range(n):
        for j in range(rowptr[i], rowptr[i + 1]):
            out.append("%d\t%d\t%f\n" % (i + 1, int(rowind[j]) + 1, float(rowval[j])))
    with open(path, "w") as f:
        f.write("".join(out))


def write_point(path, rowptr, rowind, rowval):
    """src/io.c:1419-1462 (text)."""
    n = len(rowptr) - 1
    with open(path, "w") as f:
        for i in range(n):
            for j in range(rowptr[i], rowptr[i + 1]):
                f.write("%d\t%d\t%f\n" % (i + 1, int(rowind[j]) + 1, float(rowval[j])))
