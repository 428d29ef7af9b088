"""GPU parity of the CSR SpMV (src/blas.c:67-92, 552-606) through the C ABI.

Bar: rows handled by LDS-staged row blocks (length <= short_max) are
bitwise the reference's left-to-right sums; every row is within
1e-14 relative inf-norm error of the oracle (only the summation order of
longer rows differs).
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

TOL = 1e-14


def rel_inf(a, b):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return d / s if s > 0 else d


def irregular(n, seed, lengths=(), density=0.002, ncols=None):
    """Random CSR with uniform columns plus rows of prescribed lengths."""
    ncols = ncols or n
    rng = np.random.default_rng(seed)
    deg = rng.poisson(density * ncols, n).astype(np.int64)
    for r, L in lengths:
        deg[r] = L
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    rowind = rng.integers(0, ncols, int(rowptr[-1])).astype(np.uint32)
    rowval = rng.uniform(-1, 1, int(rowptr[-1]))
    return rowptr, rowind, rowval


def check(gb, oracle, rowptr, rowind, rowval, x, plan=None, short_max=64):
    n = rowptr.size - 1
    M = gb.Matrix(rowptr, rowind, rowval, ncols=x.size, plan=plan)
    y = M.spmv(x)
    ref = oracle.spmult(n, x.size, 1, rowptr, rowind, rowval, x)
    lens = np.diff(rowptr.astype(np.int64))
    sm = (plan or {}).get("short_max", short_max) or short_max
    short = lens <= sm
    assert np.array_equal(y[short], ref[short]), "short rows must be bitwise"
    assert rel_inf(y, ref) <= TOL
    return M, y, ref


def test_config1_laplacian2d_bitwise(gb, oracle):
    rp, ri, rv = gb.laplacian(2, 256)
    x = gb.uniform(0, 65536, stream=1)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert np.array_equal(y, ref)
    info = M.info()
    assert info["nrows_wave"] == 0 and info["nrows_block"] == 0


def test_config2_laplacian3d_full_size_bitwise(gb, oracle):
    rp, ri, rv = gb.laplacian(3, 160)
    assert rp.size - 1 == 4096000 and rp[-1] == 28518400
    x = gb.uniform(1, 4096000, stream=0)
    oracle.set_threads(8)
    try:
        M, y, ref = check(gb, oracle, rp, ri, rv, x)
    finally:
        oracle.set_threads(1)
    assert np.array_equal(y, ref)
    # closed form at full size: A 1 = number of missing neighbours
    ones = M.spmv(np.ones(4096000))
    g = np.arange(160)
    edge = (g == 0).astype(int) + (g == 159)
    missing = edge[None, None, :] + edge[None, :, None] + edge[:, None, None]
    assert np.array_equal(ones, missing.ravel().astype(float))


def test_irregular_rows_all_bins(gb, oracle):
    n = 20000
    lengths = [(5, 65), (6, 64), (7, 4096), (8, 4097), (9, 20000), (100, 300),
               (19999, 50000), (200, 0), (201, 1)]
    rp, ri, rv = irregular(n, 1, lengths)
    x = np.random.default_rng(2).uniform(-1, 1, n)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    info = M.info()
    assert info["nrows_wave"] >= 3 and info["nrows_block"] >= 3


@pytest.mark.parametrize("plan", [
    {"short_max": 1, "wave_max": 8, "window": 256, "rowcap": 1},
    {"short_max": 8, "wave_max": 64, "window": 512, "rowcap": 64},
    {"short_max": 64, "wave_max": 1 << 30, "window": 2240, "rowcap": 512},
    {"short_max": 200, "wave_max": 1000, "window": 2048, "rowcap": 300},
    {"short_max": 2303, "wave_max": 4096, "window": 1, "rowcap": 512},
    {"grid": 64},
    {"grid": 8, "xcd_remap": 1},
    {"xcd_remap": 1, "window": 700, "rowcap": 100},
])
def test_plans_same_results(gb, oracle, plan):
    n = 5000
    rp, ri, rv = irregular(n, 3, [(10, 3000), (11, 200), (12, 70), (4999, 9000)],
                           density=0.004)
    x = np.random.default_rng(4).uniform(-1, 1, n)
    check(gb, oracle, rp, ri, rv, x, plan=plan)


def test_empty_rows_and_empty_matrix(gb, oracle):
    n = 3000
    rng = np.random.default_rng(5)
    deg = np.where(rng.random(n) < 0.7, 0, rng.integers(1, 5, n))
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    ri = rng.integers(0, n, int(rp[-1])).astype(np.uint32)
    rv = rng.uniform(-1, 1, int(rp[-1]))
    x = rng.uniform(-1, 1, n)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert np.array_equal(y, ref)
    # no nonzeros at all
    rp0 = np.zeros(n + 1, dtype=np.uint32)
    M0 = gb.Matrix(rp0, np.zeros(0, np.uint32), np.zeros(0), ncols=n)
    assert np.array_equal(M0.spmv(x), np.zeros(n))


def test_single_long_cancelling_row(gb, oracle):
    # 100000 random-sign terms: the reference's own sequential rounding error
    # is ~1e-14 of |y| here, so a reordered sum cannot promise 1e-14.  The
    # exact-order plan reproduces the reference bitwise; the default tree
    # reduction stays within the textbook bound for two summation orders,
    # |y1 - y2| <= 2 (n-1) eps sum|a_ij x_j|.
    rp = np.array([0, 100000], dtype=np.uint32)
    rng = np.random.default_rng(6)
    ri = rng.integers(0, 70000, 100000).astype(np.uint32)
    rv = rng.uniform(-1, 1, 100000)
    x = rng.uniform(-1, 1, 70000)
    ref = oracle.spmult(1, 70000, 1, rp, ri, rv, x)
    exact = gb.Matrix(rp, ri, rv, ncols=70000, plan={"exact": 1}).spmv(x)
    assert np.array_equal(exact, ref)
    y = gb.Matrix(rp, ri, rv, ncols=70000).spmv(x)
    bound = 2 * (100000 - 1) * np.finfo(float).eps * np.sum(np.abs(rv * x[ri]))
    assert abs(y[0] - ref[0]) <= bound


def test_single_entry(gb):
    M = gb.Matrix(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([3.0]))
    assert M.spmv(np.array([2.0]))[0] == 6.0


@pytest.mark.parametrize("wave_max", [64, 4096, 1 << 30])
def test_exact_plan_bitwise_every_row(gb, oracle, wave_max):
    n = 20000
    lengths = [(5, 65), (6, 64), (7, 4096), (8, 4097), (9, 20000), (100, 300),
               (19999, 50000), (200, 0), (201, 1), (202, 2304), (203, 2305)]
    rp, ri, rv = irregular(n, 16, lengths)
    x = np.random.default_rng(17).uniform(-1, 1, n)
    M = gb.Matrix(rp, ri, rv, plan={"exact": 1, "wave_max": wave_max})
    assert np.array_equal(M.spmv(x), oracle.spmult(n, n, 1, rp, ri, rv, x))


def test_nonsquare_and_special_values(gb, oracle):
    rp, ri, rv = irregular(700, 7, [(3, 500)], ncols=1300, density=0.01)
    rv[::97] = 0.0
    x = np.random.default_rng(8).uniform(-1, 1, 1300)
    x[::50] = -0.0
    check(gb, oracle, rp, ri, rv, x)


def test_repeat_is_deterministic(gb):
    rp, ri, rv = irregular(30000, 9, [(1, 10000), (2, 300)])
    x = np.random.default_rng(10).uniform(-1, 1, 30000)
    M = gb.Matrix(rp, ri, rv)
    y1, y2 = M.spmv(x), M.spmv(x)
    assert np.array_equal(y1, y2)


def test_spmult_dropin_host_pointers(gb, oracle):
    rp, ri, rv = irregular(4000, 11, [(0, 5000)], ncols=4500)
    x = np.random.default_rng(12).uniform(-1, 1, 4500)
    y = gb.spmult(4000, 4500, 1, rp, ri, rv, x)
    ref = oracle.spmult(4000, 4500, 1, rp, ri, rv, x)
    assert rel_inf(y, ref) <= TOL
    # blocks: only rows inside [blocks[0], blocks[-1]) are written
    c = np.full(4000, 7.0)
    blocks = np.array([100, 900, 2000], dtype=np.uint32)
    gb.spmult(4000, 4500, 1, rp, ri, rv, x, blocks=blocks, crowval=c)
    assert np.array_equal(c[100:2000], y[100:2000])
    assert np.all(c[:100] == 7.0) and np.all(c[2000:] == 7.0)
    # nblocks == 0 computes nothing
    c2 = np.full(4000, 5.0)
    gb.spmult(4000, 4500, 1, rp, ri, rv, x, blocks=np.array([0], np.uint32), crowval=c2)
    assert np.all(c2 == 5.0)
    # a rowptr that does not start at 0: row i is entries rowptr[i]..rowptr[i+1]-1,
    # as the reference's loop reads them (the entries before rowptr[0] are never read)
    pad = 3
    rp3 = rp.astype(np.uint32) + pad
    ri3 = np.concatenate([np.full(pad, 4499, np.uint32), ri])
    rv3 = np.concatenate([np.full(pad, np.nan), rv])
    assert np.array_equal(gb.spmult(4000, 4500, 1, rp3, ri3, rv3, x), y)
    c3 = np.full(4000, 7.0)
    gb.spmult(4000, 4500, 1, rp3, ri3, rv3, x, blocks=blocks, crowval=c3)
    assert np.array_equal(c3, c)


def test_spmm_bncols3(gb, oracle):
    rp, ri, rv = irregular(3000, 13, [(4, 700)], ncols=2000)
    B = np.random.default_rng(14).uniform(-1, 1, (2000, 3))
    bcol = B.T.ravel()
    c = gb.spmult(3000, 2000, 3, rp, ri, rv, bcol)
    ref = oracle.spmult(3000, 2000, 3, rp, ri, rv, bcol)
    assert rel_inf(c, ref) <= TOL
    lens = np.diff(rp.astype(np.int64))
    short = np.repeat(lens <= 64, 3)
    assert np.array_equal(c[short], ref[short])


def test_scale_dropin(gb):
    v = np.random.default_rng(15).uniform(-1, 1, 1000)
    assert np.array_equal(gb.scale(v, 0.37), v * 0.37)


def test_invalid_inputs_rejected(gb):
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix(np.array([0, 2], np.uint32), np.array([0, 5], np.uint32), np.ones(2), ncols=3)
    assert e.value.name == "INVALIDINPUT"
    with pytest.raises(gb.GooseberryError):
        gb.Matrix(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.ones(1),
                  plan={"short_max": 2000, "window": 2000})


# ------------------------------------------------------------ column split

def _row_sum_bound(rp, ri, rv, x):
    # |fl(sum) - sum| <= (len - 1) eps sum |terms| for any summation order
    terms = np.abs(rv * x[ri.astype(np.int64)])
    lens = np.diff(rp.astype(np.int64))
    cs = np.concatenate([[0.0], np.cumsum(terms)])
    s = cs[rp[1:].astype(np.int64)] - cs[rp[:-1].astype(np.int64)]
    return 2.0 * np.maximum(lens - 1, 0) * np.finfo(np.float64).eps * s * 1.01 + 1e-300


@pytest.mark.parametrize("P", [2, 8, 13])
def test_column_split_spmv(gb, oracle, P):
    # uniform random columns, empty rows, a long row and a non-square shape
    rp, ri, rv = irregular(20000, 40 + P, [(7, 5000), (11, 0), (12, 300)], density=0.0015,
                           ncols=30011)
    x = np.random.default_rng(P).uniform(-1, 1, 30011)
    M = gb.Matrix(rp, ri, rv, ncols=x.size, plan={"csplit": P})
    assert M.info()["plan"]["csplit"] == P
    y = M.spmv(x)
    ref = oracle.spmult(20000, 30011, 1, rp, ri, rv, x)
    assert np.all(np.abs(y - ref) <= _row_sum_bound(rp, ri, rv, x))
    assert y[11] == 0.0
    # repeatable bits; back to the unsplit plan
    assert np.array_equal(M.spmv(x), y)
    M.replan(None)
    assert M.info()["plan"]["csplit"] == 0
    assert np.all(np.abs(M.spmv(x) - ref) <= _row_sum_bound(rp, ri, rv, x))


def test_column_split_more_slices_than_columns_and_spmm(gb, oracle):
    rp, ri, rv = irregular(500, 3, density=0.3, ncols=5)
    x = np.random.default_rng(4).uniform(-1, 1, 5)
    M = gb.Matrix(rp, ri, rv, ncols=5, plan={"csplit": 16})
    ref = oracle.spmult(500, 5, 1, rp, ri, rv, x)
    assert np.all(np.abs(M.spmv(x) - ref) <= _row_sum_bound(rp, ri, rv, x))
    B = np.random.default_rng(5).uniform(-1, 1, (5, 3))
    c = M.spmm(B.T.ravel(), 3) if hasattr(M, "spmm") else None
    if c is not None:
        refc = oracle.spmult(500, 5, 3, rp, ri, rv, B.T.ravel())
        assert rel_inf(c, refc) <= 1e-14


def test_column_split_rejects_exact(gb):
    rp, ri, rv = irregular(100, 5)
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix(rp, ri, rv, plan={"csplit": 4, "exact": 1})
    assert e.value.name == "INVALIDINPUT"


def test_column_split_cg_and_batch(gb, oracle):
    # CG on a split plan: iterations within 2 of the oracle, same residual;
    # the batched solver keeps using the unsplit plan of the same matrix
    M0 = gb.generate_sdd(60000, 21, 3, 27)
    rp, ri, rv = M0.download()
    M = gb.Matrix(rp, ri, rv, plan={"csplit": 8})
    b = gb.uniform(22, 60000)
    x, info = M.cgd(b, 1e-10, gb.TOL_RELATIVE)
    xo, it, _, _ = oracle.cgd(rp, ri, rv, b, 1e-10 * np.linalg.norm(b))
    assert abs(int(info.niter) - it) <= 2
    r = b - oracle.spmv(rp, ri, rv, x)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)
    B = gb.Batch(1)
    db, dx = gb.DeviceArray.from_host(b), gb.DeviceArray(60000)
    s = B.add(M, db, dx, 1e-10, gb.TOL_RELATIVE, 0, gb.PRECISION_MIXED)
    B.solve()
    assert B.info(s).status == 1
    r = b - oracle.spmv(rp, ri, rv, dx.numpy())
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)


# ---- 16-bit column offsets in the row blocks (plan wide_index = 0)

def test_index16_used_and_same_bits_as_wide(gb, oracle):
    rp, ri, rv = gb.laplacian(3, 60)
    n = rp.size - 1
    x = gb.uniform(5, n)
    Mn, yn, ref = check(gb, oracle, rp, ri, rv, x)
    Mw, yw, _ = check(gb, oracle, rp, ri, rv, x, plan={"wide_index": 1})
    assert Mn.info()["index16"] == 1 and Mn.info()["plan"]["wide_index"] == 0
    assert Mw.info()["index16"] == 0 and Mw.info()["plan"]["wide_index"] == 1
    assert np.array_equal(yn, yw) and np.array_equal(yn, ref)
    # FP32 values read the same offsets
    xf = x.astype(np.float32)
    assert np.array_equal(Mn.spmv_f32(xf), Mw.spmv_f32(xf))
    # CG: same iterations and bits
    b = gb.uniform(6, n)
    xa, ia = Mn.cgd(b, 1e-10)
    xb, ib = Mw.cgd(b, 1e-10)
    assert ia.niter == ib.niter and np.array_equal(xa, xb)


def one_per_row(n, ncols, extra):
    """Diagonal matrix plus (row, col) entries, each row sorted."""
    rows = [[i] if i < ncols else [] for i in range(n)]
    for r, c in extra:
        rows[r] = sorted(set(rows[r]) | {c})
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint32)
    ri = np.array([c for r in rows for c in r], dtype=np.uint32)
    rv = np.random.default_rng(n + len(extra)).uniform(-1, 1, ri.size)
    return rp, ri, rv


@pytest.mark.parametrize("col,fits", [(32767, True), (32768, False)])
def test_index16_positive_limit(gb, oracle, col, fits):
    # one row block starting at row 0: the offset is the column itself
    rp, ri, rv = one_per_row(4, 40000, [(0, col), (2, 7)])
    x = np.random.default_rng(1).uniform(-1, 1, 40000)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert M.info()["nblocks_stream"] == 1
    assert M.info()["index16"] == (1 if fits else 0)
    assert np.array_equal(y, ref)


@pytest.mark.parametrize("col,fits", [(39000 - 32768, True), (39000 - 32769, False)])
def test_index16_negative_limit(gb, oracle, col, fits):
    # rowcap 1: every row is its own block, so row 39000's offset is col - 39000
    rp, ri, rv = one_per_row(40000, 40000, [(39000, col)])
    x = np.random.default_rng(2).uniform(-1, 1, 40000)
    M, y, ref = check(gb, oracle, rp, ri, rv, x, plan={"rowcap": 1})
    assert M.info()["index16"] == (1 if fits else 0)
    assert np.array_equal(y, ref)


def test_index16_replan_and_scattered_columns(gb, oracle):
    # uniform columns over 200000: offsets leave int16, the plan stays wide
    rp, ri, rv = irregular(200000, 7, density=0.00002)
    x = np.random.default_rng(8).uniform(-1, 1, 200000)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert M.info()["index16"] == 0
    # a banded matrix switched between the two by replanning
    rp, ri, rv = gb.laplacian(2, 300)
    x = gb.uniform(9, rp.size - 1)
    M = gb.Matrix(rp, ri, rv, plan={"wide_index": 1})
    y0 = M.spmv(x)
    M.replan({})
    assert M.info()["index16"] == 1
    assert np.array_equal(M.spmv(x), y0)
    M.replan({"wide_index": 1, "rowcap": 64})
    assert M.info()["index16"] == 0
    assert np.array_equal(M.spmv(x), y0)


def test_index16_block_check_beyond_row_precheck(gb, oracle):
    # row 39000's column 6000 is 33000 behind the row: within what its block
    # might allow (blocks start up to 511 rows earlier), so only the block
    # check finds that its block (rows from 38912) cannot hold it
    rp, ri, rv = one_per_row(40000, 40000, [(39000, 6000)])
    x = np.random.default_rng(3).uniform(-1, 1, 40000)
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert M.info()["index16"] == 0
    assert np.array_equal(y, ref)
    # column 6200 in the same row: -32712 from the block's first row, fits
    rp, ri, rv = one_per_row(40000, 40000, [(39000, 6200)])
    M, y, ref = check(gb, oracle, rp, ri, rv, x)
    assert M.info()["index16"] == 1
    assert np.array_equal(y, ref)


# ---- irregular rows at the SpMV contract's tolerance

@pytest.mark.parametrize("plan", [None, {"csplit": 8}, {"xtile": 4096}, {"xtile": 8192},
                                  "autotune"])
def test_zipf_spmv_rel_inf(gb, oracle, plan):
    # Zipf(1.5) row degrees clipped to [2, 4096] (rows up to a few thousand
    # entries), uniform columns: every plan the autotuner can pick, the
    # 8-slice column split and the x tiles among them, within 1e-14 relative
    # inf-norm of the reference's left-to-right sums
    n = 200000
    M = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096,
                        plan=plan if isinstance(plan, dict) else None)
    if plan == "autotune":
        M.autotune(gb.PRECISION_FP64, 2)
    rp, ri, rv = M.download()
    assert np.diff(rp.astype(np.int64)).max() > 1000
    x = gb.uniform(2, n, stream=0)
    y = M.spmv(x)
    ref = oracle.spmult(n, n, 1, rp, ri, rv, x)
    assert rel_inf(y, ref) <= TOL, (M.info()["plan"], rel_inf(y, ref))


# ---- value-indexed plans (8-bit value indices into a table, CSR-VI)

def _few_values(n, seed, nvals, lengths=(), density=0.002):
    rp, ri, _ = irregular(n, seed, lengths, density)
    rng = np.random.default_rng(seed + 100)
    table = rng.uniform(-1, 1, nvals)
    rv = table[rng.integers(0, nvals, int(rp[-1]))]
    return rp, ri, rv


def test_value_index_laplacian_same_bits(gb, oracle):
    rp, ri, rv = gb.laplacian(3, 40)
    n = rp.size - 1
    x = gb.uniform(3, n)
    Mv = gb.Matrix(rp, ri, rv)
    Mw = gb.Matrix(rp, ri, rv, plan={"wide_values": 1})
    assert Mv.info()["values8"] == 2 and Mw.info()["values8"] == 0
    yv, yw = Mv.spmv(x), Mw.spmv(x)
    assert np.array_equal(yv, yw) and np.array_equal(yv, oracle.spmv(rp, ri, rv, x))
    xf = x.astype(np.float32)
    assert np.array_equal(Mv.spmv_f32(xf), Mw.spmv_f32(xf))
    # CG: the value-indexed plan runs the pipelined kernel, whose p.Ap partials
    # are grouped by its workgroups rather than by row block, so the dot
    # products (not the SpMV) round differently
    b = gb.uniform(4, n)
    xv, iv = Mv.cgd(b, 1e-10, gb.TOL_RELATIVE)
    xw, iw = Mw.cgd(b, 1e-10, gb.TOL_RELATIVE)
    assert abs(int(iv.niter) - int(iw.niter)) <= 1
    assert np.max(np.abs(xv - xw)) <= 1e-9 * np.max(np.abs(xw))


@pytest.mark.parametrize("nvals,expect", [(1, 1), (200, 200), (256, 256), (257, 0)])
def test_value_index_table_sizes_and_long_rows(gb, oracle, nvals, expect):
    # long rows take the wavefront and workgroup paths, which read the
    # indexed values too
    rp, ri, rv = _few_values(30000, 7, nvals, lengths=[(5, 3000), (6, 9000), (7, 0)])
    x = np.random.default_rng(2).uniform(-1, 1, 30000)
    M = gb.Matrix(rp, ri, rv)
    Mw = gb.Matrix(rp, ri, rv, plan={"wide_values": 1})
    assert M.info()["values8"] == expect
    y, yw = M.spmv(x), Mw.spmv(x)
    assert np.array_equal(y, yw)
    assert rel_inf(y, oracle.spmv(rp, ri, rv, x)) <= TOL
    Me = gb.Matrix(rp, ri, rv, plan={"exact": 1})
    assert np.array_equal(Me.spmv(x), oracle.spmv(rp, ri, rv, x))


def test_value_index_signed_zero_nan_and_replans(gb, oracle):
    rp, ri, _ = irregular(5000, 9, density=0.004)
    nnz = int(rp[-1])
    rng = np.random.default_rng(10)
    rv = np.where(rng.uniform(0, 1, nnz) < 0.5, 0.0, -0.0)
    rv[::7] = 2.5
    x = rng.uniform(-1, 1, 5000)
    x[3] = np.inf  # 0 * inf = nan, -0 * inf = nan: the same in either plan
    M = gb.Matrix(rp, ri, rv)
    assert M.info()["values8"] == 3
    Mw = gb.Matrix(rp, ri, rv, plan={"wide_values": 1})
    y, yw = M.spmv(x), Mw.spmv(x)
    assert np.array_equal(y.view(np.uint64), yw.view(np.uint64))
    # a column split drops the index, the default plan brings it back
    M.replan({"csplit": 4})
    assert M.info()["values8"] == 0
    M.replan(None)
    assert M.info()["values8"] == 3
    assert np.array_equal(M.spmv(x).view(np.uint64), yw.view(np.uint64))


def test_value_index_random_values_not_indexed(gb):
    M = gb.generate_sdd(20000, 5, 3, 27)
    assert M.info()["values8"] == 0


# ---- persistent pipelined row blocks (plan grid on matrices of short rows)

@pytest.mark.parametrize("grid", [1, 7, 1024, 65536])
@pytest.mark.parametrize("wide_values", [0, 1])
def test_pipelined_row_blocks_same_bits(gb, oracle, grid, wide_values):
    rp, ri, rv = gb.laplacian(3, 36)
    n = rp.size - 1
    x = gb.uniform(6, n)
    ref = gb.Matrix(rp, ri, rv)
    M = gb.Matrix(rp, ri, rv, plan={"grid": grid, "wide_values": wide_values})
    y = M.spmv(x)
    assert np.array_equal(y, ref.spmv(x)) and np.array_equal(y, oracle.spmv(rp, ri, rv, x))
    xf = x.astype(np.float32)
    assert np.array_equal(M.spmv_f32(xf), ref.spmv_f32(xf))
    B = np.random.default_rng(8).uniform(-1, 1, (n, 3))
    assert np.array_equal(M.spmm(B.T.ravel(), 3), ref.spmm(B.T.ravel(), 3))
    b = gb.uniform(7, n)
    x1, i1 = M.cgd(b, 1e-10, gb.TOL_RELATIVE)
    xr, it, _, _ = oracle.cgd(rp, ri, rv, b, 1e-10 * np.linalg.norm(b))
    assert i1.status == 1 and abs(int(i1.niter) - it) <= 2
    assert np.linalg.norm(b - oracle.spmv(rp, ri, rv, x1)) <= 1.05e-10 * np.linalg.norm(b)
