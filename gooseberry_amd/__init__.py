"""gooseberry_amd -- Python view of libgooseberry (the MI355X sparse-solve
library) through its C ABI (include/gooseberry.h).

The product is the C library and the `gooseberry` program; this module loads
the in-tree libgooseberry.so with ctypes and mirrors the reference's
interface for the sparse-solve path:

  spmult(...)          gooseberry_spmult (include/gooseberry.h:174-184)
  cgd(...)             the reference's cgd() (src/cgd.c:40-129), host arrays
  Matrix               a device-resident CSR matrix (SpMV, SpMM, CG)
  read_*/write_*       the .csr/.mtx/grid readers and writers (src/io.c)
  countingsort_kv, permute_sparse, permute_dense, read_partition

There is no CPU fallback: every function raises if the library is missing
or the GPU call fails.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libgooseberry.so")
CLI_PATH = os.path.join(_PKG, "bin", "gooseberry")

SUCCESS = 1
ERRORS = {
    1: "SUCCESS", 2: "UNIMPLEMENTED", 3: "INVALIDINPUT", 4: "READ",
    5: "WRITE", 6: "FILENOTFOUND", 7: "PERMISSIONDENIED", 8: "UNKNOWN",
    9: "DEVICE", 10: "NOMEMORY", 11: "BREAKDOWN",
}
ERROR_CODES = {v: k for k, v in ERRORS.items()}

FORMAT_RAW = 0
FORMAT_GRID = 1
FORMAT_CSR = 2
FORMAT_POINT = 3
FORMAT_SVM = 4
FORMAT_GRAPH = 5
FORMAT_CLU = 6
FORMAT_MATRIXMARKET = 7
FORMAT_BCSR = 8

TOL_ABSOLUTE = 0
TOL_RELATIVE = 1

FAULT_ALLOC = 0
FAULT_LAUNCH = 1

REQUEST_CANCELLED = 4
REQUEST_TIMEDOUT = 5
REQUEST_REJECTED = 6

# every extern "C" symbol declared in include/gooseberry.h
EXPORTED = [
    "gooseberry_scale", "gooseberry_spmult", "gooseberry_read_sparse_matrix",
    "gooseberry_read_dense_matrix", "gooseberry_write_dense_matrix",
    "gooseberry_write_sparse_matrix", "gooseberry_read_labels",
    "gooseberry_read_partition", "gooseberry_free", "gooseberry_strerror",
    "gooseberry_last_device_error", "gooseberry_countingsort_kv",
    "gooseberry_permute_sparse", "gooseberry_permute_dense",
    "gooseberry_device_alloc", "gooseberry_device_free",
    "gooseberry_memcpy_htod", "gooseberry_memcpy_dtoh", "gooseberry_memcpy_dtod",
    "gooseberry_device_synchronize", "gooseberry_device_memory",
    "gooseberry_event_create", "gooseberry_event_destroy",
    "gooseberry_event_record", "gooseberry_event_elapsed_ms",
    "gooseberry_matrix_create", "gooseberry_matrix_create_device",
    "gooseberry_matrix_free", "gooseberry_matrix_replan",
    "gooseberry_matrix_info", "gooseberry_matrix_spmv",
    "gooseberry_matrix_spmm", "gooseberry_matrix_cgd", "gooseberry_cgd",
    "gooseberry_generate_laplacian", "gooseberry_generate_uniform",
    "gooseberry_batch_create", "gooseberry_batch_free", "gooseberry_batch_add",
    "gooseberry_batch_remove", "gooseberry_batch_run", "gooseberry_batch_solve",
    "gooseberry_batch_info", "gooseberry_batch_workspace_bytes",
    "gooseberry_server_create", "gooseberry_server_free", "gooseberry_server_submit_sdd",
    "gooseberry_server_submit", "gooseberry_server_progress", "gooseberry_server_status",
    "gooseberry_server_take_result", "gooseberry_matrix_generate_sdd",
    "gooseberry_matrix_download", "gooseberry_generate_uniform_device",
    "gooseberry_matrix_generate_spd", "gooseberry_matrix_autotune",
    "gooseberry_add_scalar", "gooseberry_matrix_pagerank", "gooseberry_pagerank",
    "gooseberry_spmultsp", "gooseberry_spmultsp_sp", "gooseberry_matrix_spmultsp",
    "gooseberry_matrix_spmultsp_dense", "gooseberry_read_sparse_matrix_device",
    "gooseberry_parse_csr_device", "gooseberry_matrix_read",
    "gooseberry_permute_cuthillmckee", "gooseberry_matrix_cuthillmckee",
    "gooseberry_matrix_spmv_f32", "gooseberry_matrix_cgd_mixed", "gooseberry_spmult_f32",
    "gooseberry_matrix_cgd_noalloc", "gooseberry_cgd_noalloc_max_bytes",
    "gooseberry_inject_fault", "gooseberry_server_cancel", "gooseberry_server_set_timeout",
    "gooseberry_scale_i64", "gooseberry_add_scalar_i64", "gooseberry_spmult_i64",
    "gooseberry_spmultsp_i64", "gooseberry_spmultsp_sp_i64",
    "gooseberry_read_sparse_matrix_i64", "gooseberry_write_sparse_matrix_i64",
    "gooseberry_matrix_create_i64", "gooseberry_matrix_create_device_i64",
    "gooseberry_cgd_i64", "gooseberry_matrix_parts",
]

PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED_JACOBI = 2
DEGREE_UNIFORM = 0
DEGREE_ZIPF = 1


class GooseberryError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        self.name = ERRORS.get(code, str(code))
        msg = "%s failed: GOOSEBERRY_ERROR_%s" % (what, self.name)
        try:
            dev = lib().gooseberry_last_device_error().decode()
            if dev and code == 9:
                msg += " (" + dev + ")"
        except Exception:
            pass
        super().__init__(msg)


class _Plan(C.Structure):
    _fields_ = [("short_max", C.c_uint32), ("wave_max", C.c_uint32),
                ("window", C.c_uint32), ("rowcap", C.c_uint32),
                ("exact", C.c_uint32), ("grid", C.c_uint32),
                ("xcd_remap", C.c_uint32), ("csplit", C.c_uint32),
                ("wide_index", C.c_uint32), ("wide_values", C.c_uint32),
                ("xtile", C.c_uint32)]


class _MatrixInfo(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("ncols", C.c_uint32),
                ("nnz", C.c_uint64), ("plan", _Plan),
                ("nblocks_stream", C.c_uint32), ("nrows_wave", C.c_uint32),
                ("nrows_block", C.c_uint32), ("hist", C.c_uint64 * 33),
                ("index16", C.c_uint32), ("values8", C.c_uint32)]


class CgdInfo(C.Structure):
    _fields_ = [("niter", C.c_uint64), ("rnorm", C.c_double),
                ("bnorm", C.c_double), ("converged", C.c_int),
                ("status", C.c_int), ("nrefine", C.c_uint64),
                ("true_rel_residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PagerankInfo(C.Structure):
    _fields_ = [("niter", C.c_uint64), ("rmse", C.c_double), ("nsinks", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("niter", C.c_uint64),
                ("nrefine", C.c_uint64), ("rel_residual", C.c_double),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("n", C.c_uint32), ("nnz", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None
_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_F64P = C.POINTER(C.c_double)


def _sig(lib):
    u32, u64, i32, f64, sz = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_size_t
    s = {
        "gooseberry_scale": (i32, [u32, _P, f64]),
        "gooseberry_spmult": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P, u32]),
        "gooseberry_matrix_spmultsp": (i32, [_P, _P, _P, C.POINTER(_P)]),
        "gooseberry_read_sparse_matrix_device": (i32, [i32, C.c_char_p, _U32P, _U32P,
                                                       C.POINTER(_P), C.POINTER(_P),
                                                       C.POINTER(_P)]),
        "gooseberry_parse_csr_device": (i32, [_P, sz, _U32P, _U32P, C.POINTER(_P),
                                              C.POINTER(_P), C.POINTER(_P)]),
        "gooseberry_matrix_read": (i32, [i32, C.c_char_p, _P, C.POINTER(_P)]),
        "gooseberry_permute_cuthillmckee": (i32, [u32, u32, _P, _P, _P, _P, u32, _P]),
        "gooseberry_matrix_cuthillmckee": (i32, [_P, _P, u32, _P]),
        "gooseberry_matrix_spmv_f32": (i32, [_P, _P, _P]),
        "gooseberry_matrix_cgd_mixed": (i32, [_P, _P, _P, f64, i32, sz, C.POINTER(CgdInfo)]),
        "gooseberry_spmult_f32": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P, u32]),
        "gooseberry_matrix_spmultsp_dense": (i32, [_P, _P, _P]),
        "gooseberry_spmultsp": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P, _P, _P, u32]),
        "gooseberry_spmultsp_sp": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P,
                                         C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                         _P, u32]),
        "gooseberry_read_sparse_matrix": (i32, [i32, C.c_char_p, _U32P, _U32P,
                                                C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
        "gooseberry_read_dense_matrix": (i32, [i32, C.c_char_p, _U32P, _U32P, C.POINTER(_P)]),
        "gooseberry_write_dense_matrix": (i32, [i32, C.c_char_p, u32, u32, _P]),
        "gooseberry_write_sparse_matrix": (i32, [i32, C.c_char_p, u32, u32, _P, _P, _P]),
        "gooseberry_read_labels": (i32, [C.c_char_p, _U32P, C.POINTER(_P)]),
        "gooseberry_read_partition": (i32, [C.c_char_p, _U32P, _U32P, C.POINTER(_P),
                                            C.POINTER(_P), C.POINTER(_P)]),
        "gooseberry_free": (None, [_P]),
        "gooseberry_strerror": (C.c_char_p, [i32]),
        "gooseberry_last_device_error": (C.c_char_p, []),
        "gooseberry_countingsort_kv": (i32, [_P, _P, u32, u32, sz, _P, _P]),
        "gooseberry_permute_sparse": (i32, [u32, u32, _P, _P, _P, _P, _P]),
        "gooseberry_permute_dense": (i32, [u32, u32, _P, _P, _P, i32]),
        "gooseberry_device_alloc": (i32, [C.POINTER(_P), sz]),
        "gooseberry_device_free": (i32, [_P]),
        "gooseberry_memcpy_htod": (i32, [_P, _P, sz]),
        "gooseberry_memcpy_dtoh": (i32, [_P, _P, sz]),
        "gooseberry_memcpy_dtod": (i32, [_P, _P, sz]),
        "gooseberry_device_synchronize": (i32, []),
        "gooseberry_device_memory": (i32, [C.POINTER(sz), C.POINTER(sz)]),
        "gooseberry_event_create": (i32, [C.POINTER(_P)]),
        "gooseberry_event_destroy": (i32, [_P]),
        "gooseberry_event_record": (i32, [_P]),
        "gooseberry_event_elapsed_ms": (i32, [_P, _P, C.POINTER(C.c_float)]),
        "gooseberry_matrix_create": (i32, [u32, u32, _P, _P, _P, C.POINTER(_Plan), C.POINTER(_P)]),
        "gooseberry_matrix_create_device": (i32, [u32, u32, _P, _P, _P, C.POINTER(_Plan),
                                                  C.POINTER(_P)]),
        "gooseberry_matrix_free": (i32, [_P]),
        "gooseberry_matrix_replan": (i32, [_P, C.POINTER(_Plan)]),
        "gooseberry_matrix_info": (i32, [_P, C.POINTER(_MatrixInfo)]),
        "gooseberry_matrix_spmv": (i32, [_P, _P, _P]),
        "gooseberry_matrix_spmm": (i32, [_P, u32, _P, _P]),
        "gooseberry_matrix_cgd": (i32, [_P, _P, _P, f64, i32, sz, _P, C.POINTER(CgdInfo)]),
        "gooseberry_matrix_cgd_noalloc": (i32, [_P, _P, _P, f64, i32, sz, C.POINTER(CgdInfo)]),
        "gooseberry_cgd_noalloc_max_bytes": (sz, []),
        "gooseberry_inject_fault": (i32, [i32, C.c_longlong]),
        "gooseberry_server_cancel": (i32, [_P, u64]),
        "gooseberry_server_set_timeout": (i32, [_P, f64]),
        "gooseberry_cgd": (i32, [u32, u32, _P, _P, _P, _P, _P, f64, sz, C.POINTER(CgdInfo)]),
        "gooseberry_generate_laplacian": (i32, [i32, u32, u32, u32, _U32P, C.POINTER(_P),
                                                C.POINTER(_P), C.POINTER(_P)]),
        "gooseberry_generate_uniform": (i32, [u64, u64, sz, f64, f64, _P]),
        "gooseberry_batch_create": (i32, [u32, f64, C.POINTER(_P)]),
        "gooseberry_batch_free": (i32, [_P]),
        "gooseberry_batch_add": (i32, [_P, _P, _P, _P, f64, i32, sz, i32, C.POINTER(u32)]),
        "gooseberry_batch_remove": (i32, [_P, u32]),
        "gooseberry_batch_run": (i32, [_P, u32, C.POINTER(u32)]),
        "gooseberry_batch_solve": (i32, [_P]),
        "gooseberry_batch_info": (i32, [_P, u32, C.POINTER(CgdInfo)]),
        "gooseberry_batch_workspace_bytes": (sz, [_P, i32]),
        "gooseberry_server_create": (i32, [sz, u32, i32, f64, i32, C.POINTER(_P)]),
        "gooseberry_server_free": (i32, [_P]),
        "gooseberry_server_submit_sdd": (i32, [_P, u64, u32, u64, u32, u32, f64]),
        "gooseberry_server_submit": (i32, [_P, u64, _P, _P, _P, f64]),
        "gooseberry_server_progress": (i32, [_P, u32, C.POINTER(RequestResult), u32,
                                             C.POINTER(u32)]),
        "gooseberry_server_status": (i32, [_P, C.POINTER(u32), C.POINTER(u32),
                                           C.POINTER(sz)]),
        "gooseberry_server_take_result": (i32, [_P, u64, _P, sz]),
        "gooseberry_matrix_generate_sdd": (i32, [u32, u64, u32, u32, C.POINTER(_Plan),
                                                 C.POINTER(_P)]),
        "gooseberry_matrix_generate_spd": (i32, [u32, u64, C.c_int, f64, u32, u32,
                                                 C.POINTER(_Plan), C.POINTER(_P)]),
        "gooseberry_matrix_download": (i32, [_P, _P, _P, _P]),
        "gooseberry_matrix_autotune": (i32, [_P, C.c_int, u32, C.POINTER(_Plan),
                                             C.POINTER(f64)]),
        "gooseberry_add_scalar": (i32, [u32, _P, f64]),
        "gooseberry_matrix_pagerank": (i32, [_P, f64, sz, f64, _P,
                                             C.POINTER(PagerankInfo)]),
        "gooseberry_pagerank": (i32, [u32, _P, _P, _P, f64, sz, f64, _P,
                                      C.POINTER(PagerankInfo)]),
        "gooseberry_generate_uniform_device": (i32, [u64, u64, sz, f64, f64, _P]),
        "gooseberry_scale_i64": (i32, [u64, _P, f64]),
        "gooseberry_add_scalar_i64": (i32, [u64, _P, f64]),
        "gooseberry_spmult_i64": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P, u32]),
        "gooseberry_spmultsp_i64": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P, _P, _P,
                                          u32]),
        "gooseberry_spmultsp_sp_i64": (i32, [u32, u32, u32, _P, _P, _P, _P, _P, _P,
                                             C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                             _P, u32]),
        "gooseberry_read_sparse_matrix_i64": (i32, [i32, C.c_char_p, _U32P, _U32P,
                                                    C.POINTER(_P), C.POINTER(_P),
                                                    C.POINTER(_P)]),
        "gooseberry_write_sparse_matrix_i64": (i32, [i32, C.c_char_p, u32, u32, _P, _P,
                                                     _P]),
        "gooseberry_matrix_create_i64": (i32, [u32, u32, _P, _P, _P, C.POINTER(_Plan),
                                               C.POINTER(_P)]),
        "gooseberry_matrix_create_device_i64": (i32, [u32, u32, _P, _P, _P,
                                                      C.POINTER(_Plan), C.POINTER(_P)]),
        "gooseberry_cgd_i64": (i32, [u32, u32, _P, _P, _P, _P, _P, f64, sz,
                                     C.POINTER(CgdInfo)]),
        "gooseberry_matrix_parts": (i32, [_P, _U32P]),
    }
    for name, (res, args) in s.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded libgooseberry (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgooseberry.so is not built (%s); run "
                               "python -m gooseberry_amd.build" % LIB_PATH)
        _LIB = C.CDLL(LIB_PATH)
        _sig(_LIB)
    return _LIB


def inject_fault(kind, after):
    """Test hook: the (after+1)-th device allocation (FAULT_ALLOC) or kernel
    launch check (FAULT_LAUNCH) from now on fails once; after < 0 disarms."""
    check(lib().gooseberry_inject_fault(kind, int(after)), "gooseberry_inject_fault")


def check(rv, what):
    if rv != SUCCESS:
        raise GooseberryError(rv, what)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _u64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint64)


# ------------------------------------------------------------ device memory

class DeviceArray:
    """A device (HBM) buffer holding a 1-D numpy-typed array."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = _P()
        check(lib().gooseberry_device_alloc(C.byref(p), max(self.nbytes, 8)),
              "gooseberry_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().gooseberry_memcpy_htod(self.ptr, _ptr(a), self.nbytes),
              "gooseberry_memcpy_htod")

    def copy_from(self, other):
        """Device-to-device copy of other (same size) into this buffer."""
        assert other.nbytes == self.nbytes
        check(lib().gooseberry_memcpy_dtod(self.ptr, other.ptr, self.nbytes),
              "gooseberry_memcpy_dtod")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().gooseberry_memcpy_dtoh(_ptr(out), self.ptr, self.nbytes),
              "gooseberry_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().gooseberry_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().gooseberry_device_synchronize(), "gooseberry_device_synchronize")


def device_memory():
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().gooseberry_device_memory(C.byref(f), C.byref(t)),
          "gooseberry_device_memory")
    return f.value, t.value


class Event:
    """A HIP event recorded on the library stream."""

    def __init__(self):
        p = _P()
        check(lib().gooseberry_event_create(C.byref(p)), "gooseberry_event_create")
        self.ptr = p.value

    def record(self):
        check(lib().gooseberry_event_record(self.ptr), "gooseberry_event_record")

    def elapsed_ms(self, stop):
        ms = C.c_float()
        check(lib().gooseberry_event_elapsed_ms(self.ptr, stop.ptr, C.byref(ms)),
              "gooseberry_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            lib().gooseberry_event_destroy(self.ptr)
        except Exception:
            pass


# ---------------------------------------------------------------- matrices

def _plan(plan):
    if plan is None:
        return None
    p = _Plan(plan.get("short_max", 0), plan.get("wave_max", 0),
              plan.get("window", 0), plan.get("rowcap", 0),
              1 if plan.get("exact") else 0, plan.get("grid", 0),
              int(plan.get("xcd_remap", 0)), plan.get("csplit", 0),
              1 if plan.get("wide_index") else 0, 1 if plan.get("wide_values") else 0,
              int(plan.get("xtile", 0)))
    return C.byref(p)


class Matrix:
    """A CSR matrix resident in HBM with its SpMV plan."""

    def __init__(self, rowptr, rowind, rowval, ncols=None, plan=None):
        rowptr, rowind, rowval = _u32(rowptr), _u32(rowind), _f64(rowval)
        nrows = rowptr.size - 1
        if ncols is None:
            ncols = nrows
        h = _P()
        check(lib().gooseberry_matrix_create(nrows, ncols, _ptr(rowptr), _ptr(rowind),
                                             _ptr(rowval), _plan(plan), C.byref(h)),
              "gooseberry_matrix_create")
        self.h = h.value
        self.nrows, self.ncols = nrows, ncols
        self.nnz = int(rowptr[-1])

    @classmethod
    def from_rowptr64(cls, rowptr, rowind, rowval, ncols=None, plan=None):
        """gooseberry_matrix_create_i64: host arrays with 64-bit row pointers
        (held as row parts above GOOSEBERRY_PART_NNZ or 2^31 entries)."""
        rowptr, rowind, rowval = _u64(rowptr), _u32(rowind), _f64(rowval)
        nrows = rowptr.size - 1
        h = _P()
        check(lib().gooseberry_matrix_create_i64(nrows, nrows if ncols is None else ncols,
                                                 _ptr(rowptr), _ptr(rowind), _ptr(rowval),
                                                 _plan(plan), C.byref(h)),
              "gooseberry_matrix_create_i64")
        return cls._from_handle(h.value)

    @classmethod
    def from_device64(cls, nrows, ncols, d_rowptr, d_rowind, d_rowval, plan=None):
        """gooseberry_matrix_create_device_i64 on device addresses (copied)."""
        h = _P()
        check(lib().gooseberry_matrix_create_device_i64(nrows, ncols, d_rowptr, d_rowind,
                                                        d_rowval, _plan(plan), C.byref(h)),
              "gooseberry_matrix_create_device_i64")
        return cls._from_handle(h.value)

    def parts(self):
        """Row parts the matrix is held as (0: one ordinary matrix)."""
        n = C.c_uint32()
        check(lib().gooseberry_matrix_parts(self.h, C.byref(n)), "gooseberry_matrix_parts")
        return n.value

    @classmethod
    def read(cls, path, fmt=FORMAT_CSR, plan=None):
        """gooseberry_matrix_read: the file parsed on the GPU, resident."""
        h = _P()
        check(lib().gooseberry_matrix_read(fmt, path.encode(), _plan(plan), C.byref(h)),
              "gooseberry_matrix_read")
        return cls._from_handle(h.value)

    @classmethod
    def _from_handle(cls, h):
        self = cls.__new__(cls)
        self.h = h
        i = self.info()
        self.nrows, self.ncols, self.nnz = i["nrows"], i["ncols"], i["nnz"]
        return self

    def download(self):
        """(rowptr, rowind, rowval) copied back from HBM."""
        rp = np.empty(self.nrows + 1, dtype=np.uint32)
        ri = np.empty(max(self.nnz, 1), dtype=np.uint32)
        rv = np.empty(max(self.nnz, 1), dtype=np.float64)
        check(lib().gooseberry_matrix_download(self.h, _ptr(rp), _ptr(ri), _ptr(rv)),
              "gooseberry_matrix_download")
        return rp, ri[:self.nnz], rv[:self.nnz]

    def info(self):
        i = _MatrixInfo()
        check(lib().gooseberry_matrix_info(self.h, C.byref(i)), "gooseberry_matrix_info")
        return {"nrows": i.nrows, "ncols": i.ncols, "nnz": i.nnz,
                "plan": {"short_max": i.plan.short_max, "wave_max": i.plan.wave_max,
                         "window": i.plan.window, "rowcap": i.plan.rowcap,
                         "exact": i.plan.exact, "grid": i.plan.grid,
                         "xcd_remap": i.plan.xcd_remap, "csplit": i.plan.csplit,
                         "wide_index": i.plan.wide_index,
                         "wide_values": i.plan.wide_values,
                         "xtile": i.plan.xtile},
                "nblocks_stream": i.nblocks_stream, "nrows_wave": i.nrows_wave,
                "nrows_block": i.nrows_block, "hist": list(i.hist),
                "index16": i.index16, "values8": i.values8}

    def replan(self, plan=None):
        check(lib().gooseberry_matrix_replan(self.h, _plan(plan)), "gooseberry_matrix_replan")

    def pagerank(self, damping=0.85, niter=0, minerror=0.0):
        """`gooseberry pagerank` on this matrix: (vector, iterations, rmse,
        number of sinks)."""
        d = DeviceArray(self.nrows)
        info = PagerankInfo()
        check(lib().gooseberry_matrix_pagerank(self.h, float(damping), int(niter),
                                               float(minerror), d.ptr, C.byref(info)),
              "gooseberry_matrix_pagerank")
        return d.numpy(), int(info.niter), float(info.rmse), int(info.nsinks)

    def pagerank_device(self, d_out, damping=0.85, niter=0, minerror=0.0):
        """gooseberry_matrix_pagerank into a device vector: (iterations, rmse,
        number of sinks)."""
        info = PagerankInfo()
        check(lib().gooseberry_matrix_pagerank(self.h, float(damping), int(niter),
                                               float(minerror), d_out.ptr, C.byref(info)),
              "gooseberry_matrix_pagerank")
        return int(info.niter), float(info.rmse), int(info.nsinks)

    def autotune(self, precision=PRECISION_FP64, reps=5):
        """Time candidate plans from the row-length histogram, keep the
        fastest; returns (plan dict, SpMV ms)."""
        p = _Plan()
        ms = C.c_double()
        check(lib().gooseberry_matrix_autotune(self.h, precision, reps, C.byref(p),
                                               C.byref(ms)), "gooseberry_matrix_autotune")
        return {k: getattr(p, k) for k, _ in _Plan._fields_}, ms.value

    def spmv_device(self, d_x, d_y):
        check(lib().gooseberry_matrix_spmv(self.h, d_x.ptr, d_y.ptr), "gooseberry_matrix_spmv")

    def spmv(self, x):
        dx = DeviceArray.from_host(x, np.float64)
        dy = DeviceArray(self.nrows)
        self.spmv_device(dx, dy)
        return dy.numpy()

    def spmm(self, bcol, bncols):
        """C (row-major nrows x bncols) = A B, B column-major (ncols x bncols)."""
        db = DeviceArray.from_host(np.ascontiguousarray(bcol, dtype=np.float64).ravel())
        dc = DeviceArray(self.nrows * bncols)
        check(lib().gooseberry_matrix_spmm(self.h, bncols, db.ptr, dc.ptr),
              "gooseberry_matrix_spmm")
        return dc.numpy()

    def cuthillmckee(self, blocks=None):
        """gooseberry_matrix_cuthillmckee, copied back to the host."""
        d = DeviceArray(self.nrows, np.uint32)
        b = _u32(blocks)
        check(lib().gooseberry_matrix_cuthillmckee(self.h, _ptr(b),
                                                   0 if b is None else b.size - 1, d.ptr),
              "gooseberry_matrix_cuthillmckee")
        return d.numpy()

    def spmultsp(self, other, plan=None):
        """C = self @ other as a new resident Matrix (rows list their columns
        in the order of their first product)."""
        h = _P()
        check(lib().gooseberry_matrix_spmultsp(self.h, other.h, _plan(plan), C.byref(h)),
              "gooseberry_matrix_spmultsp")
        return Matrix._from_handle(h.value)

    def spmultsp_dense_device(self, other, d_c):
        check(lib().gooseberry_matrix_spmultsp_dense(self.h, other.h, d_c.ptr),
              "gooseberry_matrix_spmultsp_dense")

    def cgd_device(self, d_b, d_x, tol, tol_mode=TOL_RELATIVE, niter=0, log=None):
        info = CgdInfo()
        rv = lib().gooseberry_matrix_cgd(self.h, d_b.ptr, d_x.ptr, float(tol), tol_mode,
                                         int(niter), log, C.byref(info))
        if rv not in (SUCCESS, ERROR_CODES["BREAKDOWN"]):
            check(rv, "gooseberry_matrix_cgd")
        return info

    def spmv_f32(self, x):
        """gooseberry_matrix_spmv_f32 on a host float32 vector."""
        dx = DeviceArray.from_host(np.asarray(x, dtype=np.float32), np.float32)
        dy = DeviceArray(self.nrows, np.float32)
        check(lib().gooseberry_matrix_spmv_f32(self.h, dx.ptr, dy.ptr),
              "gooseberry_matrix_spmv_f32")
        return dy.numpy()

    def cgd_mixed(self, b, tol, tol_mode=TOL_RELATIVE, niter=0):
        """gooseberry_matrix_cgd_mixed: (x, info)."""
        db = DeviceArray.from_host(b, np.float64)
        dx = DeviceArray(self.nrows)
        info = CgdInfo()
        check(lib().gooseberry_matrix_cgd_mixed(self.h, db.ptr, dx.ptr, float(tol), tol_mode,
                                                int(niter), C.byref(info)),
              "gooseberry_matrix_cgd_mixed")
        return dx.numpy(), info

    def cgd(self, b, tol, tol_mode=TOL_RELATIVE, niter=0):
        db = DeviceArray.from_host(b, np.float64)
        dx = DeviceArray(self.nrows)
        info = self.cgd_device(db, dx, tol, tol_mode, niter)
        return dx.numpy(), info

    def free(self):
        # refused (INVALIDINPUT) while a batch or server slot holds the matrix
        if getattr(self, "h", None):
            check(lib().gooseberry_matrix_free(self.h), "gooseberry_matrix_free")
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------- reference entry points

def spmult(anrows, ancols, bncols, rowptr, rowind, rowval, bcolval, blocks=None,
           crowval=None):
    """gooseberry_spmult with host arrays; returns crowval (row-major)."""
    rowptr, rowind, rowval = _u32(rowptr), _u32(rowind), _f64(rowval)
    bcolval = _f64(bcolval)
    if crowval is None:
        crowval = np.zeros(anrows * bncols, dtype=np.float64)
    blocks = _u32(blocks)
    nblocks = 0 if blocks is None else blocks.size - 1
    check(lib().gooseberry_spmult(anrows, ancols, bncols, _ptr(rowptr), _ptr(rowind),
                                  _ptr(rowval), _ptr(bcolval), _ptr(crowval),
                                  _ptr(blocks), nblocks), "gooseberry_spmult")
    return crowval


def spmult_i64(anrows, ancols, bncols, rowptr, rowind, rowval, bcolval, blocks=None,
               crowval=None):
    """gooseberry_spmult_i64: gooseberry_spmult with 64-bit row pointers."""
    rowptr, rowind, rowval = _u64(rowptr), _u32(rowind), _f64(rowval)
    bcolval = _f64(bcolval)
    if crowval is None:
        crowval = np.zeros(anrows * bncols, dtype=np.float64)
    blocks = _u32(blocks)
    nblocks = 0 if blocks is None else blocks.size - 1
    check(lib().gooseberry_spmult_i64(anrows, ancols, bncols, _ptr(rowptr), _ptr(rowind),
                                      _ptr(rowval), _ptr(bcolval), _ptr(crowval),
                                      _ptr(blocks), nblocks), "gooseberry_spmult_i64")
    return crowval


def cgd_i64(rowptr, rowind, rowval, b, error, niter=0):
    """gooseberry_cgd_i64: cgd() with 64-bit row pointers."""
    rowptr, rowind, rowval, b = _u64(rowptr), _u32(rowind), _f64(rowval), _f64(b)
    n = rowptr.size - 1
    x = np.zeros(n, dtype=np.float64)
    info = CgdInfo()
    rv = lib().gooseberry_cgd_i64(n, n, _ptr(rowptr), _ptr(rowind), _ptr(rowval), _ptr(b),
                                  _ptr(x), float(error), int(niter), C.byref(info))
    if rv not in (SUCCESS, ERROR_CODES["BREAKDOWN"]):
        check(rv, "gooseberry_cgd_i64")
    return x, info


def spmultsp(anrows, ancols, bncols, a, b, blocks=None, crowval=None):
    """gooseberry_spmultsp: a and b are (rowptr, rowind, rowval) CSR triples
    (b has ancols rows); returns the dense row-major product."""
    arp, ari, arv = _u32(a[0]), _u32(a[1]), _f64(a[2])
    brp, bri, brv = _u32(b[0]), _u32(b[1]), _f64(b[2])
    if crowval is None:
        crowval = np.zeros(anrows * bncols, dtype=np.float64)
    blocks = _u32(blocks)
    nblocks = 0 if blocks is None else blocks.size - 1
    check(lib().gooseberry_spmultsp(anrows, ancols, bncols, _ptr(arp), _ptr(ari), _ptr(arv),
                                    _ptr(brp), _ptr(bri), _ptr(brv), _ptr(crowval),
                                    _ptr(blocks), nblocks), "gooseberry_spmultsp")
    return crowval


def spmultsp_sp(anrows, ancols, bncols, a, b, blocks=None):
    """gooseberry_spmultsp_sp: the sparse product as (rowptr, rowind, rowval),
    each row's columns in the order of their first product."""
    arp, ari, arv = _u32(a[0]), _u32(a[1]), _f64(a[2])
    brp, bri, brv = _u32(b[0]), _u32(b[1]), _f64(b[2])
    blocks = _u32(blocks)
    nblocks = 0 if blocks is None else blocks.size - 1
    rp, ri, rv = _P(), _P(), _P()
    check(lib().gooseberry_spmultsp_sp(anrows, ancols, bncols, _ptr(arp), _ptr(ari),
                                       _ptr(arv), _ptr(brp), _ptr(bri), _ptr(brv),
                                       C.byref(rp), C.byref(ri), C.byref(rv),
                                       _ptr(blocks), nblocks), "gooseberry_spmultsp_sp")
    rowptr = _take(rp, anrows + 1, np.uint32)
    nnz = int(rowptr[-1])
    return rowptr, _take(ri, nnz, np.uint32), _take(rv, nnz, np.float64)


def spmult_f32(anrows, ancols, bncols, rowptr, rowind, rowval, bcolval, blocks=None,
               crowval=None):
    """gooseberry_spmult_f32 (the reference's float build) on host arrays."""
    rowptr, rowind = _u32(rowptr), _u32(rowind)
    rowval = np.ascontiguousarray(rowval, dtype=np.float32)
    bcolval = np.ascontiguousarray(bcolval, dtype=np.float32)
    if crowval is None:
        crowval = np.zeros(anrows * bncols, dtype=np.float32)
    blocks = _u32(blocks)
    nblocks = 0 if blocks is None else blocks.size - 1
    check(lib().gooseberry_spmult_f32(anrows, ancols, bncols, _ptr(rowptr), _ptr(rowind),
                                      _ptr(rowval), _ptr(bcolval), _ptr(crowval),
                                      _ptr(blocks), nblocks), "gooseberry_spmult_f32")
    return crowval


def scale(val, s):
    val = _f64(val).copy()
    check(lib().gooseberry_scale(val.size, _ptr(val), float(s)), "gooseberry_scale")
    return val


def add_scalar(vals, s):
    """gooseberry_add_scalar on a host array (returns a new array)."""
    out = np.array(vals, dtype=np.float64, copy=True)
    check(lib().gooseberry_add_scalar(out.size, _ptr(out), float(s)), "gooseberry_add_scalar")
    return out


def pagerank(rowptr, rowind, rowval, damping=0.85, niter=0, minerror=0.0):
    """gooseberry_pagerank on host CSR arrays: (vector, iterations, rmse)."""
    rowptr, rowind, rowval = _u32(rowptr), _u32(rowind), _f64(rowval)
    n = rowptr.size - 1
    out = np.zeros(n, dtype=np.float64)
    info = PagerankInfo()
    check(lib().gooseberry_pagerank(n, _ptr(rowptr), _ptr(rowind), _ptr(rowval),
                                    float(damping), int(niter), float(minerror), _ptr(out),
                                    C.byref(info)), "gooseberry_pagerank")
    return out, int(info.niter), float(info.rmse)


def cgd(rowptr, rowind, rowval, b, error, niter=0):
    """The reference's cgd() on host arrays: absolute threshold `error`."""
    rowptr, rowind, rowval, b = _u32(rowptr), _u32(rowind), _f64(rowval), _f64(b)
    n = rowptr.size - 1
    x = np.zeros(n, dtype=np.float64)
    info = CgdInfo()
    rv = lib().gooseberry_cgd(n, n, _ptr(rowptr), _ptr(rowind), _ptr(rowval), _ptr(b),
                              _ptr(x), float(error), int(niter), C.byref(info))
    if rv not in (SUCCESS, ERROR_CODES["BREAKDOWN"]):
        check(rv, "gooseberry_cgd")
    return x, info


def countingsort_kv(keys, kmin, kmax, vals=None, want_dist=True):
    keys = _u32(keys)
    vals = _u32(vals)
    out = np.empty(keys.size, dtype=np.uint32)
    dist = np.empty(int(kmax) - int(kmin) + 2, dtype=np.uint32) if want_dist else None
    check(lib().gooseberry_countingsort_kv(_ptr(keys), _ptr(vals), kmin, kmax, keys.size,
                                           _ptr(out), _ptr(dist)),
          "gooseberry_countingsort_kv")
    return out, dist


def permute_sparse(rowptr, rowind, rowval, ncols, rowperm=None, colperm=None):
    rowptr, rowind, rowval = _u32(rowptr).copy(), _u32(rowind).copy(), _f64(rowval).copy()
    check(lib().gooseberry_permute_sparse(rowptr.size - 1, ncols, _ptr(rowptr), _ptr(rowind),
                                          _ptr(rowval), _ptr(_u32(rowperm)),
                                          _ptr(_u32(colperm))),
          "gooseberry_permute_sparse")
    return rowptr, rowind, rowval


def cuthillmckee(rowptr, rowind, ncols=None, blocks=None, perm=None):
    """gooseberry_permute_cuthillmckee: perm[p] = the row placed at p."""
    rowptr, rowind = _u32(rowptr), _u32(rowind)
    n = rowptr.size - 1
    out = np.zeros(n, dtype=np.uint32) if perm is None else _u32(perm).copy()
    b = _u32(blocks)
    check(lib().gooseberry_permute_cuthillmckee(n, n if ncols is None else ncols, _ptr(rowptr),
                                                _ptr(rowind), None, _ptr(b),
                                                0 if b is None else b.size - 1, _ptr(out)),
          "gooseberry_permute_cuthillmckee")
    return out


def permute_dense(vals, nrows, ncols, rowperm=None, colperm=None, rev=False):
    vals = _f64(vals).copy()
    check(lib().gooseberry_permute_dense(nrows, ncols, _ptr(vals), _ptr(_u32(rowperm)),
                                         _ptr(_u32(colperm)), 1 if rev else 0),
          "gooseberry_permute_dense")
    return vals


def _take(p, n, dtype):
    if n == 0:
        lib().gooseberry_free(p)
        return np.zeros(0, dtype=dtype)
    ct = C.c_uint32 if dtype == np.uint32 else C.c_double
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()
    lib().gooseberry_free(p)
    return arr


def read_sparse(path, fmt=FORMAT_CSR):
    nr, nc = C.c_uint32(), C.c_uint32()
    rp, ri, rv = _P(), _P(), _P()
    check(lib().gooseberry_read_sparse_matrix(fmt, path.encode(), C.byref(nr), C.byref(nc),
                                              C.byref(rp), C.byref(ri), C.byref(rv)),
          "gooseberry_read_sparse_matrix")
    rowptr = _take(rp, nr.value + 1, np.uint32)
    nnz = int(rowptr[-1])
    return nr.value, nc.value, rowptr, _take(ri, nnz, np.uint32), _take(rv, nnz, np.float64)


def read_dense(path):
    nr, nc = C.c_uint32(), C.c_uint32()
    rv = _P()
    check(lib().gooseberry_read_dense_matrix(FORMAT_GRID, path.encode(), C.byref(nr),
                                             C.byref(nc), C.byref(rv)),
          "gooseberry_read_dense_matrix")
    return nr.value, nc.value, _take(rv, nr.value * nc.value, np.float64)


def write_dense(path, vals, nrows, ncols):
    vals = _f64(vals)
    check(lib().gooseberry_write_dense_matrix(FORMAT_GRID, path.encode(), nrows, ncols,
                                              _ptr(vals)), "gooseberry_write_dense_matrix")


def write_csr(path, rowptr, rowind, rowval, ncols, fmt=FORMAT_CSR):
    rowptr, rowind, rowval = _u32(rowptr), _u32(rowind), _f64(rowval)
    check(lib().gooseberry_write_sparse_matrix(fmt, path.encode(), rowptr.size - 1,
                                               ncols, _ptr(rowptr), _ptr(rowind),
                                               _ptr(rowval)),
          "gooseberry_write_sparse_matrix")


def _take_device(nr, rp, ri, rv):
    """Device CSR arrays returned by the library, copied to numpy and freed."""
    rowptr = np.empty(nr + 1, dtype=np.uint32)
    check(lib().gooseberry_memcpy_dtoh(_ptr(rowptr), rp, rowptr.nbytes), "dtoh")
    nnz = int(rowptr[-1])
    rowind = np.empty(nnz, dtype=np.uint32)
    rowval = np.empty(nnz, dtype=np.float64)
    if nnz:
        check(lib().gooseberry_memcpy_dtoh(_ptr(rowind), ri, rowind.nbytes), "dtoh")
        check(lib().gooseberry_memcpy_dtoh(_ptr(rowval), rv, rowval.nbytes), "dtoh")
    for p in (rp, ri, rv):
        lib().gooseberry_device_free(p)
    return rowptr, rowind, rowval


def read_sparse_device(path, fmt=FORMAT_CSR):
    """gooseberry_read_sparse_matrix_device: parsed on the GPU, returned as
    (nrows, ncols, rowptr, rowind, rowval) host copies."""
    nr, nc = C.c_uint32(), C.c_uint32()
    rp, ri, rv = _P(), _P(), _P()
    check(lib().gooseberry_read_sparse_matrix_device(fmt, path.encode(), C.byref(nr),
                                                     C.byref(nc), C.byref(rp), C.byref(ri),
                                                     C.byref(rv)),
          "gooseberry_read_sparse_matrix_device")
    return (nr.value, nc.value) + _take_device(nr.value, rp.value, ri.value, rv.value)


def parse_csr_device(text):
    """gooseberry_parse_csr_device on bytes uploaded to the device."""
    buf = np.frombuffer(bytes(text), dtype=np.uint8)
    d = DeviceArray(buf.size, np.uint8)
    if buf.size:
        d.upload(buf)
    nr, nc = C.c_uint32(), C.c_uint32()
    rp, ri, rv = _P(), _P(), _P()
    check(lib().gooseberry_parse_csr_device(d.ptr, buf.size, C.byref(nr), C.byref(nc),
                                            C.byref(rp), C.byref(ri), C.byref(rv)),
          "gooseberry_parse_csr_device")
    d.free()
    return (nr.value, nc.value) + _take_device(nr.value, rp.value, ri.value, rv.value)


def read_labels(path, nrows=0):
    n = C.c_uint32(nrows)
    p = _P()
    check(lib().gooseberry_read_labels(path.encode(), C.byref(n), C.byref(p)),
          "gooseberry_read_labels")
    return _take(p, n.value, np.uint32)


def read_partition(path, nrows=0):
    n, k = C.c_uint32(nrows), C.c_uint32()
    mp, pp, dp = _P(), _P(), _P()
    check(lib().gooseberry_read_partition(path.encode(), C.byref(n), C.byref(k),
                                          C.byref(mp), C.byref(pp), C.byref(dp)),
          "gooseberry_read_partition")
    return (_take(mp, n.value, np.uint32), _take(pp, n.value, np.uint32),
            _take(dp, k.value + 1, np.uint32))


# ------------------------------------------------------------ generators

def laplacian(dims, nx, ny=None, nz=None):
    """2D 5-point / 3D 7-point Dirichlet Laplacian as (rowptr, rowind, rowval)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = C.c_uint32()
    rp, ri, rv = _P(), _P(), _P()
    check(lib().gooseberry_generate_laplacian(dims, nx, ny, nz, C.byref(n), C.byref(rp),
                                              C.byref(ri), C.byref(rv)),
          "gooseberry_generate_laplacian")
    rowptr = _take(rp, n.value + 1, np.uint32)
    nnz = int(rowptr[-1])
    return rowptr, _take(ri, nnz, np.uint32), _take(rv, nnz, np.float64)


def uniform(seed, n, lo=-1.0, hi=1.0, stream=0):
    out = np.empty(n, dtype=np.float64)
    check(lib().gooseberry_generate_uniform(seed, stream, n, lo, hi, _ptr(out)),
          "gooseberry_generate_uniform")
    return out


def generate_sdd(n, seed, dmin=3, dmax=27, plan=None):
    """Random symmetric diagonally dominant SPD matrix generated in HBM,
    degrees uniform in [dmin, dmax] before symmetrising."""
    h = _P()
    check(lib().gooseberry_matrix_generate_sdd(n, seed, dmin, dmax, _plan(plan), C.byref(h)),
          "gooseberry_matrix_generate_sdd")
    return Matrix._from_handle(h.value)


def generate_spd(n, seed, law=DEGREE_UNIFORM, zipf_s=1.5, dmin=2, dmax=4096, plan=None):
    """Random symmetric diagonally dominant SPD matrix generated in HBM with
    degrees drawn by `law` (DEGREE_UNIFORM or DEGREE_ZIPF) before
    symmetrising."""
    h = _P()
    check(lib().gooseberry_matrix_generate_spd(n, seed, law, zipf_s, dmin, dmax,
                                                 _plan(plan), C.byref(h)),
          "gooseberry_matrix_generate_spd")
    return Matrix._from_handle(h.value)


def autotune_all(mats, precision=PRECISION_MIXED, reps=5, min_nnz=1 << 20):
    """Autotune many matrices: one timed search per class of matrices with
    the same occupied row-length buckets and nnz within a factor of 4 (the
    largest of the class is tuned), its plan applied to the whole class.
    Classes below min_nnz keep their default plans: their SpMV is
    launch-latency bound, and timings that short cannot rank plans."""
    groups = {}
    for M in mats:
        inf = M.info()
        if int(inf["nnz"]) < min_nnz:
            continue
        occupied = tuple(int(b) for b in np.nonzero(np.asarray(inf["hist"]))[0])
        key = (occupied, int(np.log2(max(int(inf["nnz"]), 1))) // 2)
        groups.setdefault(key, []).append(M)
    plans = {}
    for key, ms in groups.items():
        rep = max(ms, key=lambda m: int(m.info()["nnz"]))
        plan, t = rep.autotune(precision, reps)
        for M in ms:
            if M is not rep:
                M.replan(plan)
        plans[str(key)] = {"plan": plan, "spmv_ms": t, "matrices": len(ms)}
    return {"classes": len(groups), "plans": plans}


def uniform_device(seed, n, lo=-1.0, hi=1.0, stream=0):
    d = DeviceArray(n)
    check(lib().gooseberry_generate_uniform_device(seed, stream, n, lo, hi, d.ptr),
          "gooseberry_generate_uniform_device")
    return d


# --------------------------------------------------------- batched solves

class Batch:
    """Independent CG solves advanced together (gooseberry_batch_*)."""

    def __init__(self, max_systems, delta=0.0):
        h = _P()
        check(lib().gooseberry_batch_create(max_systems, float(delta), C.byref(h)),
              "gooseberry_batch_create")
        self.h = h.value
        self._keep = {}

    def add(self, mat, d_b, d_x, tol, tol_mode=TOL_RELATIVE, niter=0,
            precision=PRECISION_FP64):
        slot = C.c_uint32()
        check(lib().gooseberry_batch_add(self.h, mat.h, d_b.ptr, d_x.ptr, float(tol), tol_mode,
                                         int(niter), precision, C.byref(slot)),
              "gooseberry_batch_add")
        self._keep[slot.value] = (mat, d_b, d_x)
        return slot.value

    def remove(self, slot):
        check(lib().gooseberry_batch_remove(self.h, slot), "gooseberry_batch_remove")
        self._keep.pop(slot, None)

    def run(self, steps):
        running = C.c_uint32()
        check(lib().gooseberry_batch_run(self.h, steps, C.byref(running)), "gooseberry_batch_run")
        return running.value

    def solve(self):
        check(lib().gooseberry_batch_solve(self.h), "gooseberry_batch_solve")

    def info(self, slot):
        i = CgdInfo()
        check(lib().gooseberry_batch_info(self.h, slot, C.byref(i)), "gooseberry_batch_info")
        return i

    def free(self):
        if getattr(self, "h", None):
            lib().gooseberry_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching solve server (gooseberry_server_*)."""

    def __init__(self, mem_budget=0, max_active=256, precision=PRECISION_MIXED, delta=0.0,
                 keep_results=False):
        h = _P()
        check(lib().gooseberry_server_create(int(mem_budget), max_active, precision, float(delta),
                                             1 if keep_results else 0, C.byref(h)),
              "gooseberry_server_create")
        self.h = h.value
        self._keep = {}

    def submit_sdd(self, rid, n, seed, dmin=3, dmax=27, tol=1e-10):
        check(lib().gooseberry_server_submit_sdd(self.h, rid, n, seed, dmin, dmax, float(tol)),
              "gooseberry_server_submit_sdd")

    def submit(self, rid, mat, d_b, d_x, tol=1e-10):
        check(lib().gooseberry_server_submit(self.h, rid, mat.h, d_b.ptr, d_x.ptr, float(tol)),
              "gooseberry_server_submit")
        self._keep[rid] = (mat, d_b, d_x)

    def progress(self, steps, max_done=4096):
        buf = (RequestResult * max_done)()
        nd = C.c_uint32()
        check(lib().gooseberry_server_progress(self.h, steps, buf, max_done, C.byref(nd)),
              "gooseberry_server_progress")
        out = [buf[i].as_dict() for i in range(min(nd.value, max_done))]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def cancel(self, rid):
        check(lib().gooseberry_server_cancel(self.h, rid), "gooseberry_server_cancel")
        self._keep.pop(rid, None)

    def set_timeout(self, seconds):
        check(lib().gooseberry_server_set_timeout(self.h, float(seconds)),
              "gooseberry_server_set_timeout")

    def status(self):
        a, q, m = C.c_uint32(), C.c_uint32(), C.c_size_t()
        check(lib().gooseberry_server_status(self.h, C.byref(a), C.byref(q), C.byref(m)),
              "gooseberry_server_status")
        return a.value, q.value, m.value

    def take_result(self, rid, n):
        x = np.empty(n, dtype=np.float64)
        check(lib().gooseberry_server_take_result(self.h, rid, _ptr(x), n),
              "gooseberry_server_take_result")
        return x

    def free(self):
        if getattr(self, "h", None):
            lib().gooseberry_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
