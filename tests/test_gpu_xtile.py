"""The x-tile SpMV plan (plan xtile = C) through the C ABI.

The FP64 SpMV stages x through LDS in tiles of C columns; each workgroup
owns a block of rows and adds one tile's products at a time.  A row's sum
is regrouped by tile (as with a column split), so the bar is the SpMV
contract, 1e-14 relative inf-norm against the oracle (src/blas.c:67-92),
bitwise run to run, and bitwise on integer-valued data (every product and
partial sum exact).  The structural cases: rows whose run inside one tile
spans many threads' ranges, empty rows and groups, row counts and column
counts that are not multiples of the block and tile, SpMM strides, CG's
p.Ap, the FP32 SpMV (the layout over the FP32 values), the batched
engine's steps (FP32 and FP64 CG steps, FP64 refreshes) and PageRank (the
layout over the normalised values).
"""
import numpy as np
import pytest

from test_gpu_spmv import irregular, rel_inf

pytestmark = pytest.mark.gpu


def _scattered(n, ncols, seed, lengths=(), density=0.002, integer=False):
    rp, ri, rv = irregular(n, seed, lengths, density=density, ncols=ncols)
    if integer:
        rv = np.round(rv * 8.0)
    return rp, ri, rv


@pytest.mark.parametrize("xtile", [256, 1024, 8192])
def test_xtile_exact_on_integers(gb, oracle, xtile):
    # integer values and x: every regrouped sum is exact, so the plan must
    # give the oracle's bits; row 3's 30000 entries all sit in tile 0, so its
    # run there spans about 30 entries of every thread's range
    n, ncols = 40001, 70001
    rp, ri, rv = _scattered(n, ncols, 31, [(3, 30000), (4, 5000), (40000, 2000), (10, 0)],
                            integer=True)
    k3 = slice(int(rp[3]), int(rp[4]))
    ri[k3] = np.random.default_rng(32).integers(0, min(xtile, ncols), k3.stop - k3.start)
    x = np.random.default_rng(33).integers(-8, 9, ncols).astype(np.float64)
    M = gb.Matrix(rp, ri, rv, ncols=ncols, plan={"xtile": xtile})
    assert M.info()["plan"]["xtile"] == xtile
    y = M.spmv(x)
    assert np.array_equal(y, oracle.spmult(n, ncols, 1, rp, ri, rv, x))
    assert np.array_equal(M.spmv(x), y)


@pytest.mark.parametrize("xtile", [512, 4096])
def test_xtile_random_values_vs_oracle(gb, oracle, xtile):
    n, ncols = 60000, 50000
    rp, ri, rv = _scattered(n, ncols, 34, [(0, 1500), (59999, 900)], density=0.0006)
    x = np.random.default_rng(35).uniform(-1, 1, ncols)
    ref = oracle.spmult(n, ncols, 1, rp, ri, rv, x)
    M = gb.Matrix(rp, ri, rv, ncols=ncols, plan={"xtile": xtile})
    y = M.spmv(x)
    assert rel_inf(y, ref) <= 1e-14
    # rows of at most one entry per tile are summed in the reference's order
    y2 = gb.Matrix(rp, ri, rv, ncols=ncols).spmv(x)
    assert rel_inf(y, y2) <= 1e-14
    # SpMM: the same columns with a row stride
    B = np.random.default_rng(36).uniform(-1, 1, (ncols, 3))
    C = M.spmm(B.T.ravel(), 3).reshape(n, 3)
    for j in range(3):
        assert rel_inf(C[:, j], oracle.spmult(n, ncols, 1, rp, ri, rv, B[:, j])) <= 1e-14


def test_xtile_empty_and_tiny(gb, oracle):
    # a matrix smaller than one tile and one row block per CU; empty rows
    rp, ri, rv = _scattered(300, 200, 37, [(5, 0), (6, 150)], density=0.05, integer=True)
    x = np.random.default_rng(38).integers(-4, 5, 200).astype(np.float64)
    M = gb.Matrix(rp, ri, rv, ncols=200, plan={"xtile": 256})
    assert np.array_equal(M.spmv(x), oracle.spmult(300, 200, 1, rp, ri, rv, x))
    rp0 = np.zeros(1001, np.uint32)
    M0 = gb.Matrix(rp0, np.zeros(0, np.uint32), np.zeros(0), ncols=1000, plan={"xtile": 1024})
    assert np.array_equal(M0.spmv(np.ones(1000)), np.zeros(1000))


def test_xtile_cg_and_other_paths(gb, oracle):
    # CG adds p.Ap over the whole vectors after the tiled SpMV
    M = gb.generate_spd(60000, 5, gb.DEGREE_ZIPF, 1.5, 2, 400)
    rp, ri, rv = M.download()
    n = rp.size - 1
    b = gb.uniform(6, n)
    err = 1e-10 * np.linalg.norm(b)
    xr, it, rmse, _ = oracle.cgd(rp, ri, rv, b, err)
    Mt = gb.Matrix(rp, ri, rv, plan={"xtile": 4096})
    x, info = Mt.cgd(b, err, gb.TOL_ABSOLUTE)
    assert info.converged == 1 and abs(int(info.niter) - it) <= 2, (info.niter, it)
    assert np.linalg.norm(b - oracle.spmv(rp, ri, rv, x)) <= 5e-10 * np.linalg.norm(b)
    # the FP32 SpMV on the layout: FP32 sums regrouped by tile
    Mp = gb.Matrix(rp, ri, rv)
    xf = np.random.default_rng(7).uniform(-1, 1, n).astype(np.float32)
    # each row within the bound of any FP32 summation order of its L
    # products, (L + 1) 2^-24 sum |a x|, of the FP64 sum of the same FP32
    # values
    v32 = rv.astype(np.float32).astype(np.float64)
    x64 = xf.astype(np.float64)
    y64 = oracle.spmv(rp, ri, v32, x64)
    bound = (np.diff(rp.astype(np.int64)) + 1) * 2.0 ** -24 * \
        oracle.spmv(rp, ri, np.abs(v32), np.abs(x64))
    for yf in (Mt.spmv_f32(xf), Mp.spmv_f32(xf)):
        assert np.all(np.abs(yf.astype(np.float64) - y64) <= bound)
    # PageRank runs the layout over the normalised values (same quotients,
    # sums regrouped by tile)
    pt, pp = Mt.pagerank(niter=20)[0], Mp.pagerank(niter=20)[0]
    assert np.max(np.abs(pt - pp)) <= 1e-12 * np.max(np.abs(pp))
    # replans in and out of the layout
    Mt.replan({"csplit": 4})
    assert Mt.info()["plan"]["xtile"] == 0
    Mt.replan({"xtile": 8192})
    xv = gb.uniform(8, n)
    assert rel_inf(Mt.spmv(xv), oracle.spmv(rp, ri, rv, xv)) <= 1e-14


@pytest.mark.parametrize("plan", [{"xtile": 300}, {"xtile": 128}, {"xtile": 16384},
                                  {"xtile": 1024, "csplit": 4}, {"xtile": 1024, "exact": 1}])
def test_xtile_refused_plans(gb, plan):
    rp, ri, rv = gb.laplacian(2, 32)
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix(rp, ri, rv, plan=plan)
    assert e.value.name == "INVALIDINPUT"


def test_autotune_fp64_then_fp32(gb, oracle):
    # x of 2.4 MB: the tuner tries column splits and x tiles for the FP64
    # SpMV; tuning the FP32 SpMV next starts from that plan without x tiles
    M = gb.generate_sdd(300000, 9, 3, 27)
    rp, ri, rv = M.download()
    x = gb.uniform(10, 300000)
    plan64, _ = M.autotune(gb.PRECISION_FP64, 2)
    assert rel_inf(M.spmv(x), oracle.spmv(rp, ri, rv, x)) <= 1e-14
    plan32, _ = M.autotune(gb.PRECISION_MIXED, 2)
    assert plan32["xtile"] == 0 or plan32["csplit"] == 0
    assert rel_inf(M.spmv(x), oracle.spmv(rp, ri, rv, x)) <= 1e-14
    M.replan(dict(plan64))
    assert M.info()["plan"]["xtile"] == plan64["xtile"]


def test_xtile_f32_spmv_exact_on_integers(gb, oracle):
    # FP32 values and x of small integers: every FP32 product and sum is
    # exact, so the layout's FP32 SpMV gives the oracle's values
    n, ncols = 40001, 70001
    rp, ri, rv = _scattered(n, ncols, 41, [(3, 30000), (4, 5000), (40000, 2000)],
                            integer=True)
    k3 = slice(int(rp[3]), int(rp[4]))
    ri[k3] = np.random.default_rng(42).integers(0, 8192, k3.stop - k3.start)
    x = np.random.default_rng(43).integers(-8, 9, ncols).astype(np.float32)
    ref = oracle.spmult(n, ncols, 1, rp, ri, rv, x.astype(np.float64))
    for xtile in (256, 8192):
        M = gb.Matrix(rp, ri, rv, ncols=ncols, plan={"xtile": xtile})
        y = M.spmv_f32(x)
        assert np.array_equal(y.astype(np.float64), ref), xtile


def _batch(gb, systems, precision, tol=1e-10):
    B = gb.Batch(len(systems))
    bufs, slots = [], []
    for M, b in systems:
        db = gb.DeviceArray.from_host(b)
        dx = gb.DeviceArray(M.nrows)
        slots.append(B.add(M, db, dx, tol, gb.TOL_RELATIVE, 0, precision))
        bufs.append((db, dx))
    B.solve()
    out = [(dx.numpy(), B.info(s)) for (_, dx), s in zip(bufs, slots)]
    B.free()
    return out


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_xtile_batched_solves(gb, oracle, precision):
    # a Zipf SPD system on x tiles in the batched engine: its CG steps
    # (FP32 or FP64) and FP64 refreshes run on the layout; the residual is
    # met, the iterations stay near the row-block plan's, and the system
    # gives the same bits alone and beside a row-block and a column-split
    # system
    M0 = gb.generate_spd(60000, 5, gb.DEGREE_ZIPF, 1.5, 2, 400)
    rp, ri, rv = M0.download()
    n = rp.size - 1
    b = gb.uniform(6, n)
    Mt = gb.Matrix(rp, ri, rv, plan={"xtile": 4096})
    (x_alone, i_alone), = _batch(gb, [(Mt, b)], precision)
    assert i_alone.status == 1, i_alone.as_dict()
    r = b - oracle.spmv(rp, ri, rv, x_alone)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)
    # iterations near the row-block plan's: the FP32 inner recurrence rounds
    # with the SpMV's summation order (regrouped by tile and lane here), and
    # a count of 16-20 Jacobi-scaled inner iterations moves by a few with it
    (x_r, i_r), = _batch(gb, [(M0, b)], precision)
    assert abs(int(i_r.niter) - int(i_alone.niter)) <= max(5, 0.25 * int(i_r.niter))
    rng = np.random.default_rng(9)
    L = gb.Matrix(*gb.laplacian(2, 40))
    S0 = gb.generate_sdd(30000, 32)
    S = gb.Matrix(*S0.download(), plan={"csplit": 4})
    mix = [(L, rng.uniform(-1, 1, L.nrows)), (Mt, b), (S, rng.uniform(-1, 1, 30000))]
    res = _batch(gb, mix, precision)
    assert res[1][1].niter == i_alone.niter and res[1][1].nrefine == i_alone.nrefine
    assert np.array_equal(res[1][0], x_alone)
    (x_l, i_l), = _batch(gb, [mix[0]], precision)
    assert np.array_equal(res[0][0], x_l) and res[0][1].niter == i_l.niter
    for x, info in res:
        assert info.status == 1
