"""Explicitly typed FP32 entry points and the one-system mixed-precision
solve.

gooseberry_spmult_f32 is the reference's float build of
src/blas.c:67-92 (real_t = float): each row sums val*x left to right from
+0.0f, every product and sum rounded to single precision.  numpy float32
arithmetic rounds the same way, so the expected result is built one entry
position at a time across all rows (a sequential sum per row); rows of up
to 64 entries must match it bit for bit."""
import numpy as np
import pytest

from test_oracle_spgemm import random_csr

pytestmark = pytest.mark.gpu


def seq_f32(rp, ri, rv, x):
    """Row-sequential float32 sums: y[i] = ((0 + v0 x0) + v1 x1) + ..."""
    rp = rp.astype(np.int64)
    n = rp.size - 1
    lens = np.diff(rp)
    y = np.zeros(n, dtype=np.float32)
    v = rv.astype(np.float32)
    x = x.astype(np.float32)
    for k in range(int(lens.max()) if n else 0):
        rows = np.flatnonzero(lens > k)
        idx = rp[rows] + k
        y[rows] = y[rows] + v[idx] * x[ri[idx]]
    return y


def test_spmult_f32_bitwise(gb):
    rp, ri, rv = gb.laplacian(3, 30)
    n = rp.size - 1
    rng = np.random.default_rng(50)
    rv = rng.uniform(-1, 1, rv.size).astype(np.float32)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = gb.spmult_f32(n, n, 1, rp, ri, rv, x)
    assert np.array_equal(y.view(np.uint32), seq_f32(rp, ri, rv, x).view(np.uint32))
    # irregular rows (up to 64 entries), three right-hand sides, blocks
    a = random_csr(3000, 2500, 12, 51, empty_every=7, dups=True)
    arp, ari = a[0], a[1]
    assert np.diff(arp.astype(np.int64)).max() <= 64
    arv = rng.uniform(-1, 1, ari.size).astype(np.float32)
    B = rng.uniform(-1, 1, (3, 2500)).astype(np.float32)        # column-major B
    C = gb.spmult_f32(3000, 2500, 3, arp, ari, arv, B.ravel()).reshape(3000, 3)
    for j in range(3):
        assert np.array_equal(C[:, j].view(np.uint32),
                              seq_f32(arp, ari, arv, B[j]).view(np.uint32))
    sentinel = np.full(3000, 5.0, dtype=np.float32)
    part = gb.spmult_f32(3000, 2500, 1, arp, ari, arv, B[0], blocks=[100, 900, 1200],
                         crowval=sentinel.copy())
    ref = seq_f32(arp, ari, arv, B[0])
    assert np.array_equal(part[100:1200].view(np.uint32), ref[100:1200].view(np.uint32))
    assert np.all(part[:100] == 5.0) and np.all(part[1200:] == 5.0)


def test_resident_spmv_f32(gb):
    rp, ri, rv = gb.laplacian(2, 200)
    n = rp.size - 1
    x = np.random.default_rng(52).uniform(-1, 1, n).astype(np.float32)
    M = gb.Matrix(rp, ri, rv)
    y = M.spmv_f32(x)
    assert np.array_equal(y.view(np.uint32), seq_f32(rp, ri, rv, x).view(np.uint32))


def test_cgd_mixed_one_system(gb, oracle):
    rp, ri, rv = gb.laplacian(3, 40)
    n = rp.size - 1
    b = gb.uniform(3, n)
    M = gb.Matrix(rp, ri, rv)
    x, info = M.cgd_mixed(b, 1e-10, gb.TOL_RELATIVE)
    assert info.status == 1 and info.converged == 1 and info.nrefine >= 1
    r = b - oracle.spmv(rp, ri, rv, x)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
    xr, it, _, _ = oracle.cgd(rp, ri, rv, b, 1e-12 * np.linalg.norm(b), niter=5000)
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
