"""GPU test of the allocation-free CG entry point,
gooseberry_matrix_cgd_noalloc, through the C ABI: two solves in a row (the
second on a smaller system) against gooseberry_matrix_cgd, to the CG parity
bar; and its refusal of a system whose workspace exceeds the array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc(gb, M, db, dx, tol, tol_mode, niter=0):
    info = gb.CgdInfo()
    rv = gb.lib().gooseberry_matrix_cgd_noalloc(M.h, db.ptr, dx.ptr, float(tol), tol_mode,
                                               int(niter), C.byref(info))
    return rv, info


def test_noalloc_cg_matches_allocating_entry_point(gb, oracle):
    # a 3D Laplacian, then a smaller 2D one: the shared workspace keeps the
    # first call's contents, which the second must not depend on
    for dims, side, seed in ((3, 48, 1), (2, 100, 2)):
        rp, ri, rv = gb.laplacian(dims, side)
        M = gb.Matrix(rp, ri, rv)
        n = M.nrows
        b = gb.uniform(seed, n)
        db = gb.DeviceArray.from_host(b)
        dx1, dx2 = gb.DeviceArray(n), gb.DeviceArray(n)
        i1 = M.cgd_device(db, dx1, 1e-10, gb.TOL_RELATIVE)
        rv2, i2 = noalloc(gb, M, db, dx2, 1e-10, gb.TOL_RELATIVE)
        assert rv2 == gb.SUCCESS
        # the same kernels in the same order: the same bits and counts
        assert i2.status == i1.status == 1 and i2.niter == i1.niter
        assert i2.rnorm == i1.rnorm and i2.bnorm == i1.bnorm
        x1, x2 = dx1.numpy(), dx2.numpy()
        assert np.array_equal(x1, x2)
        # and the CG parity bar against the oracle
        xr, it, _, _ = oracle.cgd(rp, ri, rv, b, 1e-10 * np.linalg.norm(b))
        assert abs(int(i2.niter) - it) <= 2
        r = b - oracle.spmv(rp, ri, rv, x2)
        assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)


def test_noalloc_cg_on_x_tiles(gb, oracle):
    # a plan whose SpMV stages x through LDS in column tiles: the same
    # kernels as the allocating entry point, no workspace beyond the array
    M0 = gb.generate_spd(60000, 5, gb.DEGREE_ZIPF, 1.5, 2, 400)
    rp, ri, rv = M0.download()
    M = gb.Matrix(rp, ri, rv, plan={"xtile": 4096})
    n = M.nrows
    b = gb.uniform(4, n)
    db = gb.DeviceArray.from_host(b)
    dx1, dx2 = gb.DeviceArray(n), gb.DeviceArray(n)
    i1 = M.cgd_device(db, dx1, 1e-10, gb.TOL_RELATIVE)
    rv2, i2 = noalloc(gb, M, db, dx2, 1e-10, gb.TOL_RELATIVE)
    assert rv2 == gb.SUCCESS
    assert i2.status == i1.status == 1 and i2.niter == i1.niter
    assert np.array_equal(dx1.numpy(), dx2.numpy())
    r = b - oracle.spmv(rp, ri, rv, dx2.numpy())
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)


def test_noalloc_cg_cap_and_absolute_tolerance(gb):
    rp, ri, rv = gb.laplacian(2, 64)
    M = gb.Matrix(rp, ri, rv)
    b = gb.uniform(3, M.nrows)
    db = gb.DeviceArray.from_host(b)
    dx1, dx2 = gb.DeviceArray(M.nrows), gb.DeviceArray(M.nrows)
    err = 1e-9 * float(np.linalg.norm(b))
    i1 = M.cgd_device(db, dx1, err, gb.TOL_ABSOLUTE, niter=17)
    rv2, i2 = noalloc(gb, M, db, dx2, err, gb.TOL_ABSOLUTE, niter=17)
    assert rv2 == gb.SUCCESS and i1.status == i2.status == 2 and i2.niter == 17
    assert np.array_equal(dx1.numpy(), dx2.numpy())


def test_noalloc_cg_refuses_beyond_its_array(gb):
    # 24 bytes per row of workspace: a diagonal system just past the array's
    # capacity is refused before any launch; the library stays usable
    cap = gb.lib().gooseberry_cgd_noalloc_max_bytes()
    n = cap // 24 + 1024
    rp = np.arange(n + 1, dtype=np.uint32)
    ri = np.arange(n, dtype=np.uint32)
    M = gb.Matrix(rp, ri, np.full(n, 2.0))
    db, dx = gb.DeviceArray.from_host(np.ones(n)), gb.DeviceArray(n)
    rv2, _ = noalloc(gb, M, db, dx, 1e-10, gb.TOL_RELATIVE)
    assert rv2 == gb.ERROR_CODES["INVALIDINPUT"]
    info = M.cgd_device(db, dx, 1e-10, gb.TOL_RELATIVE)
    assert info.status == 1 and np.all(dx.numpy() == 0.5)
