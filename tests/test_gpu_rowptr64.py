"""64-bit row pointers (the reference's GOOSEBERRY_64BIT_INDICES build,
include/gooseberry.h:65-71) through the C ABI.

A matrix with more entries than one row part holds (2^31, or
GOOSEBERRY_PART_NNZ, read when a matrix is made) is held as row parts, each
with 32-bit row pointers and a plan of its own.  Every row is still summed
by the same kernel code from its own first entry, so the SpMV is bitwise the
32-bit matrix's; CG adds p.Ap over the whole vectors in another fixed order,
so it meets the oracle's bar (same relative residual, iterations within
+-2).  The small cases force parts with GOOSEBERRY_PART_NNZ; one case builds
a matrix of more than 2^32 entries in HBM.
"""
import ctypes as C
import os

import numpy as np
import pytest

from test_gpu_spmv import irregular, rel_inf

pytestmark = pytest.mark.gpu


@pytest.fixture
def part_nnz():
    old = os.environ.get("GOOSEBERRY_PART_NNZ")

    def set_(v):
        if v is None:
            os.environ.pop("GOOSEBERRY_PART_NNZ", None)
        else:
            os.environ["GOOSEBERRY_PART_NNZ"] = str(int(v))

    yield set_
    set_(old)


def test_row_parts_spmv_bitwise(gb, oracle, part_nnz):
    rp, ri, rv = irregular(30000, 21, [(5, 10000), (7000, 300), (29999, 2000)])
    x = np.random.default_rng(22).uniform(-1, 1, 30000)
    y32 = gb.Matrix(rp, ri, rv).spmv(x)
    ref = oracle.spmult(30000, 30000, 1, rp, ri, rv, x)
    assert rel_inf(y32, ref) <= 1e-14
    nnz = int(rp[-1])
    rp64 = rp.astype(np.uint64)
    # parts end on row boundaries: 4, 10 and 183 parts of at most pn entries
    for pn, want in ((nnz // 3 + 1, 4), (200000, 10), (10000, 183)):
        part_nnz(pn)
        M = gb.Matrix.from_rowptr64(rp64, ri, rv)
        assert M.parts() == want, (pn, M.parts())
        assert M.info()["nnz"] == nnz
        assert np.array_equal(M.spmv(x), y32), pn
        B = np.random.default_rng(23).uniform(-1, 1, (30000, 3))
        assert np.array_equal(M.spmm(B.T.ravel(), 3),
                              gb.Matrix(rp, ri, rv).spmm(B.T.ravel(), 3))
    # the FP32 SpMV over the parts' views of one FP32 copy: the unparted
    # matrix's bits (the same row kernels on the same rows)
    xf = x.astype(np.float32)
    y32f = gb.Matrix(rp, ri, rv).spmv_f32(xf)
    for pn in (nnz // 3 + 1, 10000):
        part_nnz(pn)
        M = gb.Matrix.from_rowptr64(rp64, ri, rv)
        assert M.parts() > 1
        assert np.array_equal(M.spmv_f32(xf), y32f), pn
        M.replan({"short_max": 32})   # new plans per part, the same copy
        assert np.array_equal(M.spmv_f32(xf), gb.Matrix(rp, ri, rv, plan={"short_max": 32})
                              .spmv_f32(xf)), pn
    # within one part: an ordinary matrix
    part_nnz(None)
    M = gb.Matrix.from_rowptr64(rp64, ri, rv)
    assert M.parts() == 0 and np.array_equal(M.spmv(x), y32)
    # a row of more entries than a part holds
    part_nnz(9999)
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix.from_rowptr64(rp64, ri, rv)
    assert e.value.name == "INVALIDINPUT"


def test_row_parts_plans(gb, part_nnz):
    # the parts of a Laplacian take the 16-bit offsets and the value table,
    # and replanning rebuilds every part's plan with the same products
    rp, ri, rv = gb.laplacian(3, 60)
    n = rp.size - 1
    x = gb.uniform(4, n, stream=1)
    y32 = gb.Matrix(rp, ri, rv).spmv(x)
    part_nnz(int(rp[-1]) // 5)
    M = gb.Matrix.from_rowptr64(rp.astype(np.uint64), ri, rv)
    assert M.parts() == 6
    info = M.info()
    assert info["index16"] == 1 and info["values8"] == 2
    assert np.array_equal(M.spmv(x), y32)
    M.replan({"wide_index": 1, "wide_values": 1})
    info = M.info()
    assert info["index16"] == 0 and info["values8"] == 0
    assert np.array_equal(M.spmv(x), y32)
    for plan in ({"csplit": 4}, {"xtile": 1024}):
        with pytest.raises(gb.GooseberryError) as e:
            M.replan(plan)
        assert e.value.name == "INVALIDINPUT"
    assert np.array_equal(M.spmv(x), y32)


def test_row_parts_cg_vs_oracle(gb, oracle, part_nnz):
    rp, ri, rv = gb.laplacian(3, 40)
    n = rp.size - 1
    b = gb.uniform(1, n)
    err = 1e-10 * np.linalg.norm(b)
    xr, it, rmse, _ = oracle.cgd(rp, ri, rv, b, err)
    part_nnz(int(rp[-1]) // 4 + 1)
    M = gb.Matrix.from_rowptr64(rp.astype(np.uint64), ri, rv)
    assert M.parts() == 4
    x, info = M.cgd(b, err, gb.TOL_ABSOLUTE)
    assert info.converged == 1 and info.rnorm <= err
    assert abs(int(info.niter) - it) <= 2, (info.niter, it)
    r = b - oracle.spmv(rp, ri, rv, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 5e-10
    assert np.max(np.abs(x - xr)) <= 1e-8 * np.max(np.abs(xr))
    # the host-array entry point with 64-bit row pointers
    x2, info2 = gb.cgd_i64(rp.astype(np.uint64), ri, rv, b, err)
    assert info2.niter == info.niter and np.array_equal(x2, x)
    # the allocation-free solve runs the same kernels on row parts
    dx = gb.DeviceArray(n)
    db = gb.DeviceArray.from_host(b)
    info3 = gb.CgdInfo()
    rv3 = gb.lib().gooseberry_matrix_cgd_noalloc(M.h, db.ptr, dx.ptr, float(err),
                                                 gb.TOL_ABSOLUTE, 0, C.byref(info3))
    assert rv3 == gb.SUCCESS
    assert info3.niter == info.niter and np.array_equal(dx.numpy(), x)


def test_spmult_i64_matches_spmult(gb, part_nnz):
    rp, ri, rv = irregular(4000, 11, [(0, 5000)], ncols=4500)
    x = np.random.default_rng(12).uniform(-1, 1, 4500)
    y = gb.spmult(4000, 4500, 1, rp, ri, rv, x)
    rp64 = rp.astype(np.uint64)
    assert np.array_equal(gb.spmult_i64(4000, 4500, 1, rp64, ri, rv, x), y)
    blocks = np.array([100, 900, 2000], dtype=np.uint32)
    c = np.full(4000, 7.0)
    gb.spmult_i64(4000, 4500, 1, rp64, ri, rv, x, blocks=blocks, crowval=c)
    assert np.array_equal(c[100:2000], y[100:2000])
    assert np.all(c[:100] == 7.0) and np.all(c[2000:] == 7.0)
    # rows held as parts inside the call
    part_nnz(int(rp[-1]) // 3)
    assert np.array_equal(gb.spmult_i64(4000, 4500, 1, rp64, ri, rv, x), y)


def test_row_parts_refuse_what_they_do_not_take(gb, part_nnz):
    rp, ri, rv = gb.laplacian(2, 64)
    part_nnz(int(rp[-1]) // 2)
    M = gb.Matrix.from_rowptr64(rp.astype(np.uint64), ri, rv)
    assert M.parts() == 2
    n = rp.size - 1
    b = gb.uniform(2, n)
    for call in (lambda: M.cgd_mixed(b, 1e-10),
                 lambda: M.autotune(gb.PRECISION_FP64, 2),
                 lambda: M.pagerank(niter=3),
                 lambda: M.download(),
                 lambda: M.cuthillmckee()):
        with pytest.raises(gb.GooseberryError) as e:
            call()
        assert e.value.name == "UNIMPLEMENTED"
    # the matrix still works after the refusals
    assert np.array_equal(M.spmv(b), gb.Matrix(rp, ri, rv).spmv(b))
    # and takes the FP32 SpMV: each part's value table rounded to FP32
    bf = b.astype(np.float32)
    assert M.info()["values8"] > 0
    assert np.array_equal(M.spmv_f32(bf), gb.Matrix(rp, ri, rv).spmv_f32(bf))


def _fill_periodic(gb, dst, chunk, total_bytes):
    """dst (a device address) gets `chunk` (host bytes) repeated to
    total_bytes: one upload, then device-to-device doubling."""
    lib = gb.lib()
    n0 = chunk.nbytes
    gb.check(lib.gooseberry_memcpy_htod(C.c_void_p(dst), chunk.ctypes.data_as(C.c_void_p),
                                        n0), "gooseberry_memcpy_htod")
    have = n0
    while have < total_bytes:
        k = min(have, total_bytes - have)
        gb.check(lib.gooseberry_memcpy_dtod(C.c_void_p(dst + have), C.c_void_p(dst), k),
                 "gooseberry_memcpy_dtod")
        have += k


def test_more_than_2_32_entries(gb):
    """2^22 rows x 1025 entries (4,299,161,600 > 2^32) built in HBM: row i
    holds columns 0..1024 with values ((i + j) mod 7) - 3.  The matrix is
    held as three row parts; every y_i is checked exactly (integer values,
    every product and sum exact in FP64), and the 7-row period of the values
    catches a row or entry offset that is off."""
    n, w = 1 << 22, 1025
    nnz = n * w
    assert nnz > (1 << 32)
    rows = 7 * 1024  # a whole number of value periods
    i = np.arange(rows, dtype=np.int64).reshape(-1, 1)
    j = np.arange(w, dtype=np.int64).reshape(1, -1)
    ind_chunk = np.ascontiguousarray(np.broadcast_to(j, (rows, w)), dtype=np.uint32).ravel()
    val_chunk = (((i + j) % 7) - 3).astype(np.float64).ravel()
    rp = gb.DeviceArray.from_host(np.arange(n + 1, dtype=np.uint64) * np.uint64(w), np.uint64)
    ri = gb.DeviceArray(nnz, np.uint32)
    rv = gb.DeviceArray(nnz, np.float64)
    _fill_periodic(gb, ri.ptr, ind_chunk, nnz * 4)
    _fill_periodic(gb, rv.ptr, val_chunk, nnz * 8)
    M = gb.Matrix.from_device64(n, n, rp.ptr, ri.ptr, rv.ptr)
    ri.free()
    rv.free()
    rp.free()
    assert M.parts() == 3 and M.info()["nnz"] == nnz
    x = (np.arange(n) % 1000).astype(np.float64)
    y = M.spmv(x)
    vals = ((np.arange(7).reshape(-1, 1) + np.arange(w).reshape(1, -1)) % 7 - 3).astype(np.float64)
    per = vals @ x[:w]  # y_i = per[i mod 7], exact
    assert np.array_equal(y, per[np.arange(n) % 7])
    M.free()


def test_c_program_with_64bit_indices(gb, tmp_path):
    """A C program built with GOOSEBERRY_64BIT_INDICES reads a .csr file and
    calls gooseberry_spmult under the reference's names with 64-bit row
    pointers: its y is the 32-bit entry point's, bit for bit."""
    import subprocess
    from test_host import build_prog64
    exe = build_prog64(gb, tmp_path)
    rp, ri, rv = irregular(3000, 31, [(2, 500), (2999, 5000)])
    a = os.path.join(str(tmp_path), "a.csr")
    b = os.path.join(str(tmp_path), "b.csr")
    gb.write_csr(a, rp, ri, rv, 3000)
    r = subprocess.run([exe, a, b, "spmult"], stdout=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.returncode
    lines = r.stdout.split("\n")
    nr, nc, nnz = map(int, lines[0].split())
    y = np.array([float(v) for v in lines[1:1 + nr]])
    nr2, nc2, rp2, ri2, rv2 = gb.read_sparse(a)
    assert (nr2, nc2, int(rp2[-1])) == (nr, nc, nnz)
    x = 1.0 + (np.arange(nc2) % 7) / 8.0
    assert np.array_equal(y, gb.spmult(nr, nc2, 1, rp2, ri2, rv2, x))
