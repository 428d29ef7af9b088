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


def cgd(rowptr, rowind, rowval, b, error, niter=0, trace=False):
    """Returns (x, iterations, final rmse, trace array (k x 3) or None)."""
    rowptr, rowind, rowval, b = _u32(rowptr), _u32(rowind), _f64(rowval), _f64(b)
    n = rowptr.size - 1
    x = np.zeros(n)
    cap = int(niter) if (trace and niter) else (20000 if trace else 0)
    tr = np.zeros(3 * cap) if cap else None
    it = C.c_size_t()
    rm = C.c_double()
    rv = lib().orc_cgd(n, n, _p(rowptr), _p(rowind), _p(rowval), n, _p(b), n, _p(x),
                       float(error), int(niter), None, _p(tr), cap, C.byref(it), C.byref(rm))
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
