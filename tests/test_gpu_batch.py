"""GPU tests of the batched solver, the mixed-precision (FP32 inner, FP64
reliable-update) CG, the generated SDD systems and the continuous-batching
server, through the C ABI.

Bars: every solve reaches its relative residual (checked with the oracle's
FP64 SpMV); a system gives bitwise the same solution and iteration count
alone and inside any mix; the mixed-precision result meets the FP64
tolerance against the FP64 oracle.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def true_rel(oracle, rp, ri, rv, x, b):
    r = b - oracle.spmv(rp, ri, rv, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def assert_symmetric_entries(rp, ri, rv):
    """The stored entries of A and of A^T are the same multiset of
    (row, column, value) triples (positions drawn more than once are kept as
    separate entries, whose sums scipy would add in different orders)."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    cols = ri.astype(np.int64)
    o1 = np.lexsort((rv, cols, rows))
    o2 = np.lexsort((rv, rows, cols))
    assert np.array_equal(rows[o1], cols[o2])
    assert np.array_equal(cols[o1], rows[o2])
    assert np.array_equal(rv[o1], rv[o2])


def lap(gb, dims, side):
    rp, ri, rv = gb.laplacian(dims, side)
    return gb.Matrix(rp, ri, rv), (rp, ri, rv)


# ------------------------------------------------------------- generator

def test_generated_sdd_matches_restatement(gb):
    # the generator restated in numpy (tests/spd_restatement.py): bitwise
    # equal CSR arrays for the uniform degree law
    from spd_restatement import generate_uniform
    for n, seed, dmin, dmax in ((5000, 7, 3, 27), (3001, 8, 2, 9), (2, 1, 1, 3)):
        M = gb.generate_sdd(n, seed, dmin, dmax)
        got = M.download()
        want = generate_uniform(n, seed, dmin, dmax)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def test_generated_sdd_structure(gb):
    n = 50000
    M = gb.generate_sdd(n, 7, 3, 27)
    rp, ri, rv = M.download()
    assert_symmetric_entries(rp, ri, rv)
    lens = np.diff(rp.astype(np.int64))
    # out-degree d - 1 in [2, 26] plus in-edges (mean 14) plus the diagonal
    assert lens.min() >= 3 and abs(lens.mean() - 29.0) < 0.2
    # diagonal last: 1.01 x the absolute off-diagonal sum in row order
    for i in np.random.default_rng(0).integers(0, n, 300):
        a, b = int(rp[i]), int(rp[i + 1])
        assert ri[b - 1] == i
        s = 0.0
        for k in range(a, b - 1):
            assert -1.0 <= rv[k] < 0.0 and ri[k] != i
            s = s + (-rv[k])
        assert rv[b - 1] == 1.01 * s
    counts = np.bincount(ri, minlength=n)
    assert np.count_nonzero(counts) == n
    M2 = gb.generate_sdd(n, 7, 3, 27)
    assert all(np.array_equal(u, v) for u, v in zip(M2.download(), (rp, ri, rv)))
    M3 = gb.generate_sdd(n, 8, 3, 27)
    assert not np.array_equal(M3.download()[1][:1000], ri[:1000])


def test_generated_zipf_degrees(gb):
    # Zipf(1.5) clipped to [2, 4096]: nnz = n + 2 sum(d_i - 1); the degree
    # draws restated in numpy agree except where the two pow() differ in the
    # last bit at an acceptance boundary
    from spd_restatement import zipf_degrees
    n = 200000
    M = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    info = M.info()
    d = zipf_degrees(n, 2, 1.5, 2, 4096)
    want = n + 2 * int((d - 1).sum())
    assert abs(int(info["nnz"]) - want) <= 2 * 4096 * 2
    rp, ri, rv = M.download()
    assert_symmetric_entries(rp, ri, rv)
    A = sp.csr_matrix((rv.copy(), ri.copy(), rp.copy()), shape=(n, n))
    diag = rv[rp[1:].astype(np.int64) - 1]
    assert np.all(ri[rp[1:].astype(np.int64) - 1] == np.arange(n))
    off = np.abs(A).sum(axis=1).A1 - diag
    assert np.allclose(diag, 1.01 * off, rtol=1e-12)


# ---------------------------------------------------------------- batches

def solve_batch(gb, systems, tol=1e-10, precision=0, niter=0, order=None):
    """systems: list of (Matrix, b) -> list of (x, info) in input order."""
    B = gb.Batch(max(len(systems), 1))
    bufs = []
    idx = list(range(len(systems))) if order is None else order
    slots = {}
    for i in idx:
        M, b = systems[i]
        db = gb.DeviceArray.from_host(b)
        dx = gb.DeviceArray(M.nrows)
        slots[i] = B.add(M, db, dx, tol, gb.TOL_RELATIVE, niter, precision)
        bufs.append((db, dx))
    B.solve()
    out = [None] * len(systems)
    for k, i in enumerate(idx):
        out[i] = (bufs[k][1].numpy(), B.info(slots[i]))
    return out


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batch_solves_converge(gb, oracle, precision):
    rng = np.random.default_rng(1)
    mats = [lap(gb, 2, 60), lap(gb, 3, 24), lap(gb, 2, 17)]
    sdd = gb.generate_sdd(30000, 3, 3, 27)
    systems = [(M, rng.uniform(-1, 1, M.nrows)) for M, _ in mats]
    systems.append((sdd, rng.uniform(-1, 1, sdd.nrows)))
    res = solve_batch(gb, systems, 1e-10, precision)
    hosts = [h for _, h in mats] + [sdd.download()]
    for (x, info), (M, b), (rp, ri, rv) in zip(res, systems, hosts):
        assert info.status == 1, info.as_dict()
        tr = true_rel(oracle, rp, ri, rv, x, b)
        assert tr <= 1.05e-10, tr
        if precision >= 1:
            assert info.nrefine >= 1 and info.true_rel_residual <= 1e-10


def test_batch_fp64_matches_single_solver_iterations(gb, oracle):
    M, (rp, ri, rv) = lap(gb, 2, 96)
    b = gb.uniform(4, M.nrows)
    (x, info), = solve_batch(gb, [(M, b)], 1e-10, 0)
    xs, info_s = M.cgd(b, 1e-10, gb.TOL_RELATIVE)
    assert abs(int(info.niter) - int(info_s.niter)) <= 2
    xr, it, _, _ = oracle.cgd(rp, ri, rv, b, 1e-10 * np.linalg.norm(b))
    assert abs(int(info.niter) - it) <= 2


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_alone_and_in_the_mix_bitwise(gb, precision):
    rng = np.random.default_rng(2)
    target, _ = lap(gb, 3, 20)
    bt = rng.uniform(-1, 1, target.nrows)
    others = [lap(gb, 2, 40)[0], gb.generate_sdd(20000, 11), lap(gb, 3, 12)[0],
              gb.generate_sdd(70000, 12)]
    (x_alone, i_alone), = solve_batch(gb, [(target, bt)], 1e-10, precision)
    mix = [(M, rng.uniform(-1, 1, M.nrows)) for M in others]
    mix.insert(2, (target, bt))
    res = solve_batch(gb, mix, 1e-10, precision, order=[4, 1, 2, 0, 3])
    x_mix, i_mix = res[2]
    assert i_mix.niter == i_alone.niter and i_mix.nrefine == i_alone.nrefine
    assert np.array_equal(x_mix, x_alone)


def test_batch_caps_and_zero_rhs(gb):
    M, _ = lap(gb, 2, 50)
    b = gb.uniform(5, M.nrows)
    res = solve_batch(gb, [(M, b), (M, np.zeros(M.nrows))], 1e-12, 0, niter=9)
    assert res[0][1].status == 2 and res[0][1].niter == 9
    assert res[1][1].status == 1 and res[1][1].niter == 0 and np.all(res[1][0] == 0)
    res = solve_batch(gb, [(M, b)], 1e-12, 1, niter=9)
    assert res[0][1].status == 2 and res[0][1].niter == 9


def test_mixed_matches_fp64_oracle_solution(gb, oracle):
    # the FP64 oracle: the reference's double CG run to 1e-12
    M, (rp, ri, rv) = lap(gb, 3, 32)
    b = gb.uniform(6, M.nrows)
    xr, _, _, _ = oracle.cgd(rp, ri, rv, b, 1e-12 * np.linalg.norm(b))
    (x, info), = solve_batch(gb, [(M, b)], 1e-10, 1)
    assert info.status == 1
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert err < 1e-8, err


# ----------------------------------------------------------------- server

def test_server_generated_requests(gb, oracle):
    S = gb.Server(max_active=8, precision=gb.PRECISION_MIXED, keep_results=True)
    sizes = [4096, 30000, 9000, 120000, 5000, 60000, 17000, 4096, 80000, 25000]
    for i, n in enumerate(sizes):
        S.submit_sdd(i, n, 1000 + i, 3, 27, 1e-10)
    done = []
    for _ in range(2000):
        done += S.progress(8)
        if len(done) == len(sizes):
            break
    assert sorted(r["id"] for r in done) == list(range(len(sizes)))
    for r in done:
        assert r["status"] == 1 and r["rel_residual"] <= 1e-10
        assert r["submit_s"] <= r["admit_s"] <= r["done_s"]
    # every solution checked against the matrix regenerated alone, and the
    # solution equal bitwise to the same request solved alone
    for rid in (1, 3, 8):
        n = sizes[rid]
        x = S.take_result(rid, n)
        M = gb.generate_sdd(n, 1000 + rid, 3, 27)
        rp, ri, rv = M.download()
        b = gb.uniform_device(1000 + rid, n, stream=1).numpy()
        assert true_rel(oracle, rp, ri, rv, x, b) <= 1.05e-10
        S1 = gb.Server(max_active=1, precision=gb.PRECISION_MIXED, keep_results=True)
        S1.submit_sdd(0, n, 1000 + rid, 3, 27, 1e-10)
        while not S1.progress(8):
            pass
        assert np.array_equal(S1.take_result(0, n), x)


def test_server_memory_admission_queues(gb):
    # a 20000-row generated FP64 request (degrees up to 27) is charged
    # ~13.8 MB while admitted (matrix upper bound, workspace, b and x) and
    # needs ~16.8 MB more while it is generated: a 36 MiB budget admits one
    # at a time and the rest wait in the queue
    n = 20000
    budget = 36 << 20
    S = gb.Server(mem_budget=1, max_active=4, precision=gb.PRECISION_FP64)
    S.submit_sdd(0, n, 1, 3, 27, 1e-8)
    with pytest.raises(gb.GooseberryError) as e:
        S.progress(1)
    assert e.value.name == "NOMEMORY"
    S.free()
    S = gb.Server(mem_budget=budget, max_active=4, precision=gb.PRECISION_FP64)
    for i in range(4):
        S.submit_sdd(i, n, 10 + i, 3, 27, 1e-8)
    peak = 0
    done = []
    for _ in range(1000):
        done += S.progress(4)
        a, q, m = S.status()
        peak = max(peak, a)
        assert m <= budget
        if len(done) == 4:
            break
    assert len(done) == 4 and peak == 1
    assert all(r["status"] == 1 for r in done)


def test_server_user_matrices(gb, oracle):
    S = gb.Server(max_active=4, precision=gb.PRECISION_FP64)
    items = []
    for i, side in enumerate((30, 45, 22)):
        M, host = lap(gb, 2, side)
        b = gb.uniform(20 + i, M.nrows)
        db, dx = gb.DeviceArray.from_host(b), gb.DeviceArray(M.nrows)
        S.submit(i, M, db, dx, 1e-9)
        items.append((M, host, b, dx))
    done = []
    while len(done) < 3:
        done += S.progress(8)
    for M, (rp, ri, rv), b, dx in items:
        assert true_rel(oracle, rp, ri, rv, dx.numpy(), b) <= 1.05e-9


# ---------------------------------------------------------------- autotune

def _row_bound(rp, ri, rv, x):
    # |fl(sum) - sum| <= (len - 1) eps sum |terms| for any summation order
    terms = np.abs(rv * x[ri.astype(np.int64)])
    lens = np.diff(rp.astype(np.int64))
    s = np.add.reduceat(terms, rp[:-1].astype(np.int64)) if terms.size else terms
    s = np.where(lens > 0, s, 0.0)
    return 2.0 * np.maximum(lens - 1, 0) * np.finfo(np.float64).eps * s


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_autotune_keeps_results(gb, oracle, precision):
    n = 30000
    M = gb.generate_spd(n, 4, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    rp, ri, rv = M.download()
    x = np.random.default_rng(3).uniform(-1, 1, n)
    yr = oracle.spmv(rp, ri, rv, x)
    plan, ms = M.autotune(precision, 3)
    assert ms > 0
    got = M.info()["plan"]
    assert all(got[k] == plan[k] for k in ("short_max", "wave_max", "window", "rowcap"))
    assert plan["window"] + plan["short_max"] <= 2304 and plan["rowcap"] <= 512
    y = M.spmv(x)
    assert np.all(np.abs(y - yr) <= _row_bound(rp, ri, rv, x))
    # a tuned matrix solves in a batch as before
    b = gb.uniform(9, n)
    (xs, info), = solve_batch(gb, [(M, b)], 1e-10, precision)
    assert info.status == 1
    assert true_rel(oracle, rp, ri, rv, xs, b) <= 1.05e-10


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_split_matrix_in_batch_alone_and_mix(gb, oracle, precision):
    # a column-split matrix runs its CG steps over its slices (refresh steps
    # over its own plan); same bits alone and in a mix, residual met
    M0 = gb.generate_sdd(80000, 31, 3, 27)
    rp, ri, rv = M0.download()
    Ms = gb.Matrix(rp, ri, rv, plan={"csplit": 6})
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, 80000)
    (x_alone, i_alone), = solve_batch(gb, [(Ms, b)], 1e-10, precision)
    assert i_alone.status == 1
    assert true_rel(oracle, rp, ri, rv, x_alone, b) <= 1.05e-10
    others = [lap(gb, 2, 40)[0], gb.generate_sdd(30000, 32)]
    mix = [(M, rng.uniform(-1, 1, M.nrows)) for M in others]
    mix.insert(1, (Ms, b))
    res = solve_batch(gb, mix, 1e-10, precision, order=[2, 1, 0])
    x_mix, i_mix = res[1]
    assert i_mix.niter == i_alone.niter and np.array_equal(x_mix, x_alone)
    # iterations close to the unsplit plan's (FP32 slice sums of a few terms
    # round less than whole-row FP32 sums, which can save mixed iterations)
    (x_u, i_u), = solve_batch(gb, [(M0, b)], 1e-10, precision)
    assert abs(int(i_u.niter) - int(i_alone.niter)) <= max(3, 0.15 * int(i_u.niter))


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batch_index16_and_wide_plans_bitwise(gb, precision):
    # the batched SpMV reads each matrix's 16-bit offsets when its plan has
    # them, 32-bit indices otherwise; the same systems give the same bits
    rng = np.random.default_rng(5)
    shapes = [(2, 50), (3, 22), (2, 31), (3, 9)]
    hosts = [gb.laplacian(d, s) for d, s in shapes]
    bs = [rng.uniform(-1, 1, rp.size - 1) for rp, _, _ in hosts]
    narrow = [gb.Matrix(*h) for h in hosts]
    wide = [gb.Matrix(*h, plan={"wide_index": 1}) for h in hosts]
    assert all(M.info()["index16"] == 1 for M in narrow)
    assert all(M.info()["index16"] == 0 for M in wide)
    # mixed: two narrow and two wide matrices in one batch
    mixed = [narrow[0], wide[1], narrow[2], wide[3]]
    ra = solve_batch(gb, list(zip(narrow, bs)), 1e-10, precision)
    rb = solve_batch(gb, list(zip(wide, bs)), 1e-10, precision)
    rc = solve_batch(gb, list(zip(mixed, bs)), 1e-10, precision)
    for (xa, ia), (xb, ib), (xc, ic) in zip(ra, rb, rc):
        assert ia.status == 1
        assert ia.niter == ib.niter == ic.niter
        assert np.array_equal(xa, xb) and np.array_equal(xa, xc)


def test_held_matrix_refuses_replan_autotune_and_free(gb):
    # a batch slot copies the plan's arrays: while it holds the matrix the
    # plan may not be rebuilt and the matrix may not be freed
    M, _ = lap(gb, 2, 40)
    n = M.nrows
    db = gb.DeviceArray.from_host(gb.uniform(2, n))
    dx = gb.DeviceArray(n)
    B = gb.Batch(2)
    slot = B.add(M, db, dx, 1e-8, gb.TOL_RELATIVE, 0, gb.PRECISION_FP64)
    for call in (lambda: M.replan({"short_max": 32}),
                 lambda: M.autotune(gb.PRECISION_FP64, 1),
                 lambda: M.free()):
        with pytest.raises(gb.GooseberryError) as e:
            call()
        assert e.value.name == "INVALIDINPUT"
    B.solve()
    assert B.info(slot).status == 1
    x1 = dx.numpy()
    # after the slot is removed, all three work again and the solve repeats
    B.remove(slot)
    M.replan({"short_max": 32})
    M.replan(None)
    dx2 = gb.DeviceArray(n)
    slot = B.add(M, db, dx2, 1e-8, gb.TOL_RELATIVE, 0, gb.PRECISION_FP64)
    B.solve()
    assert np.array_equal(dx2.numpy(), x1)
    B.free()  # releases its slots
    M.free()


def test_server_free_releases_held_matrices(gb):
    # the server's active requests hold their matrices until it is freed
    S = gb.Server(max_active=4, precision=gb.PRECISION_MIXED)
    for i in range(3):
        S.submit_sdd(i, 6000 + 100 * i, 50 + i, 3, 9, 1e-10)
    S.progress(1)
    a, _, _ = S.status()
    assert a == 3
    S.free()


# ------------------------------------------------ Jacobi-scaled mixed solves

def test_jacobi_mixed_fewer_iterations_same_answer(gb, oracle):
    # Zipf row degrees give diagonals from ~2 to a few thousand: the
    # Jacobi-scaled inner solve needs far fewer iterations than plain CG,
    # and its answer meets the FP64 tolerance against the FP64 oracle (the
    # reference's double CG run to 1e-12)
    n = 60000
    M = gb.generate_spd(n, 5, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    rp, ri, rv = M.download()
    b = gb.uniform(10, n)
    xr, it_ref, _, _ = oracle.cgd(rp, ri, rv, b, 1e-12 * np.linalg.norm(b))
    (x64, i64), (xm, im), (xj, ij) = [solve_batch(gb, [(M, b)], 1e-10, p)[0]
                                      for p in (0, 1, 2)]
    for x, info in ((x64, i64), (xm, im), (xj, ij)):
        assert info.status == 1
        assert true_rel(oracle, rp, ri, rv, x, b) <= 1.05e-10
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-8
    assert ij.true_rel_residual <= 1e-10 and ij.nrefine >= 1
    assert 3 * int(ij.niter) < int(i64.niter), (ij.niter, i64.niter)
    assert 3 * int(ij.niter) < int(im.niter), (ij.niter, im.niter)


def test_jacobi_laplacian_same_iterations_as_mixed(gb, oracle):
    # a constant diagonal: the scaled system is A / 6, so the inner
    # iterations match plain mixed precision up to FP32 rounding
    M, (rp, ri, rv) = lap(gb, 3, 30)
    b = gb.uniform(11, M.nrows)
    (xm, im), = solve_batch(gb, [(M, b)], 1e-10, 1)
    (xj, ij), = solve_batch(gb, [(M, b)], 1e-10, 2)
    assert im.status == ij.status == 1
    assert abs(int(im.niter) - int(ij.niter)) <= max(3, 0.1 * int(im.niter))
    assert true_rel(oracle, rp, ri, rv, xj, b) <= 1.05e-10


def test_jacobi_refuses_nonpositive_diagonal(gb):
    # row 1 has no diagonal entry, row 2 a negative one
    rp = np.array([0, 2, 3, 5], np.uint32)
    ri = np.array([0, 1, 0, 1, 2], np.uint32)
    rv = np.array([2.0, -1.0, -1.0, -1.0, -3.0])
    M = gb.Matrix(rp, ri, rv)
    db, dx = gb.DeviceArray.from_host(np.ones(3)), gb.DeviceArray(3)
    B = gb.Batch(1)
    with pytest.raises(gb.GooseberryError) as e:
        B.add(M, db, dx, 1e-10, gb.TOL_RELATIVE, 0, gb.PRECISION_MIXED_JACOBI)
    assert e.value.name == "INVALIDINPUT"
    # the same matrix still solves in the other precisions' paths
    B.add(M, db, dx, 1e-10, gb.TOL_RELATIVE, 5, gb.PRECISION_FP64)


def test_server_jacobi_alone_equals_mix(gb, oracle):
    S = gb.Server(max_active=6, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    sizes = [4096, 50000, 9000, 120000, 20000, 70000]
    for i, n in enumerate(sizes):
        S.submit_sdd(i, n, 2000 + i, 3, 27, 1e-10)
    done = []
    for _ in range(2000):
        done += S.progress(8)
        if len(done) == len(sizes):
            break
    assert sorted(r["id"] for r in done) == list(range(len(sizes)))
    assert all(r["status"] == 1 and r["rel_residual"] <= 1e-10 for r in done)
    for rid in (1, 3):
        n = sizes[rid]
        x = S.take_result(rid, n)
        rp, ri, rv = gb.generate_sdd(n, 2000 + rid, 3, 27).download()
        b = gb.uniform_device(2000 + rid, n, stream=1).numpy()
        assert true_rel(oracle, rp, ri, rv, x, b) <= 1.05e-10
        S1 = gb.Server(max_active=1, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
        S1.submit_sdd(0, n, 2000 + rid, 3, 27, 1e-10)
        while not S1.progress(8):
            pass
        assert np.array_equal(S1.take_result(0, n), x)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batch_value_index_same_bits(gb, precision):
    # value-indexed plans (the Laplacians have 2 distinct values, their
    # Jacobi-scaled FP32 values too) give the batched solves the same bits
    rng = np.random.default_rng(12)
    shapes = [(2, 45), (3, 20), (2, 28)]
    hosts = [gb.laplacian(d, s) for d, s in shapes]
    bs = [rng.uniform(-1, 1, rp.size - 1) for rp, _, _ in hosts]
    vi = [gb.Matrix(*h) for h in hosts]
    wide = [gb.Matrix(*h, plan={"wide_values": 1}) for h in hosts]
    assert all(M.info()["values8"] == 2 for M in vi)
    ra = solve_batch(gb, list(zip(vi, bs)), 1e-10, precision)
    rb = solve_batch(gb, list(zip(wide, bs)), 1e-10, precision)
    for (xa, ia), (xb, ib) in zip(ra, rb):
        assert ia.status == ib.status == 1 and ia.niter == ib.niter
        assert np.array_equal(xa, xb)
