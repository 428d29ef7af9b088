"""The irregular, batched and served configurations at their stated sizes,
through size-independent properties (the oracle cannot run them in
seconds):

* the 2,000,000-row Zipf(1.5) SPD system: every precision reaches relative
  residual 1e-10, checked with the FP64 SpMV, and the mixed-precision
  answers agree with the FP64 reference (plain CG run to 1e-12);
* 512 Laplacians (2D/3D alternating, n log-uniform in [1e3, 2e5]) solved in
  one batch: all converge; a sample of 2D and 3D systems checked against the
  oracle's FP64 SpMV; every result against its FP64 reference;
* a 6-second Poisson stream at 200 requests/s through the server: every
  request converges, and sampled requests give the same bits alone.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel_res(gb, M, db, dx):
    dy = gb.DeviceArray(M.nrows)
    M.spmv_device(dx, dy)
    b = db.numpy()
    return float(np.linalg.norm(b - dy.numpy()) / np.linalg.norm(b))


@pytest.mark.parametrize("plan", [None, {"xtile": 8192}])
def test_zipf_2m_every_precision(gb, plan):
    # the default plan, and the x tiles the autotuner picks for this matrix
    # (the batched engine's mixed solves then run on the layout too)
    n = 2000000
    M = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096, plan=plan)
    assert int(M.info()["nnz"]) > 300000000
    db = gb.uniform_device(2, n, stream=1)
    dref = gb.DeviceArray(n)
    iref = M.cgd_device(db, dref, 1e-12, gb.TOL_RELATIVE)
    assert iref.status == 1
    xref = dref.numpy()
    results = {}
    dx = gb.DeviceArray(n)
    info = M.cgd_device(db, dx, TOL, gb.TOL_RELATIVE)
    results["fp64"] = (info, dx.numpy())
    assert info.status == 1 and rel_res(gb, M, db, dx) <= 1.05 * TOL
    for name, prec in (("mixed", gb.PRECISION_MIXED), ("jacobi", gb.PRECISION_MIXED_JACOBI)):
        dx = gb.DeviceArray(n)
        B = gb.Batch(1)
        s = B.add(M, db, dx, TOL, gb.TOL_RELATIVE, 0, prec)
        B.solve()
        info = B.info(s)
        B.free()
        assert info.status == 1 and info.true_rel_residual <= TOL, name
        assert rel_res(gb, M, db, dx) <= 1.05 * TOL, name
        results[name] = (info, dx.numpy())
    for name, (info, x) in results.items():
        err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
        assert err < 1e-8, (name, err)
    # the scaled inner solve needs far fewer iterations than plain CG
    assert 3 * int(results["jacobi"][0].niter) < int(results["fp64"][0].niter)


def test_512_laplacians_in_one_batch(gb, oracle):
    u = gb.uniform(3, 512, 0.0, 1.0, stream=2)
    mats, hosts, rhs, xs = [], [], [], []
    for i in range(512):
        dims = 2 if i % 2 == 0 else 3
        target = np.exp(np.log(1000.0) + u[i] * (np.log(200000.0) - np.log(1000.0)))
        side = max(2, int(round(target ** (1.0 / dims))))
        h = gb.laplacian(dims, side)
        M = gb.Matrix(*h)
        mats.append(M)
        hosts.append(h)
        rhs.append(gb.DeviceArray.from_host(gb.uniform(3 + i, M.nrows, stream=1)))
        xs.append(gb.DeviceArray(M.nrows))
    assert 15e6 < sum(M.nrows for M in mats) < 25e6
    for prec in (gb.PRECISION_MIXED_JACOBI, gb.PRECISION_FP64):
        B = gb.Batch(512)
        slots = [B.add(M, b, x, TOL, gb.TOL_RELATIVE, 0, prec)
                 for M, b, x in zip(mats, rhs, xs)]
        B.solve()
        infos = [B.info(s) for s in slots]
        B.free()
        assert all(i.status == 1 for i in infos)
        # the oracle's FP64 SpMV on 2D and 3D systems, small and large
        for k in (0, 1, 2, 3, 254, 255, 510, 511):
            rp, ri, rv = hosts[k]
            b = rhs[k].numpy()
            r = b - oracle.spmv(rp, ri, rv, xs[k].numpy())
            assert np.linalg.norm(r) <= 1.05 * TOL * np.linalg.norm(b), k
        # every result
        assert max(rel_res(gb, M, b, x) for M, b, x in zip(mats, rhs, xs)) <= 1.05 * TOL
    # mixed results against the FP64 reference (plain CG to 1e-12) on every
    # 16th system of each kind
    B = gb.Batch(512)
    slots = [B.add(M, b, x, TOL, gb.TOL_RELATIVE, 0, gb.PRECISION_MIXED_JACOBI)
             for M, b, x in zip(mats, rhs, xs)]
    B.solve()
    B.free()
    for k in list(range(0, 512, 16)) + list(range(1, 512, 16)):
        dref = gb.DeviceArray(mats[k].nrows)
        mats[k].cgd_device(rhs[k], dref, 1e-12, gb.TOL_RELATIVE)
        xr = dref.numpy()
        assert np.linalg.norm(xs[k].numpy() - xr) / np.linalg.norm(xr) < 1e-6, k


def test_served_stream_6s(gb):
    rate, duration = 200.0, 6.0
    m = int(rate * duration * 1.5) + 64
    u = gb.uniform(5, 2 * m, 0.0, 1.0, stream=3)
    t_arr = np.cumsum(-np.log1p(-u[:m]) / rate)
    t_arr = t_arr[t_arr < duration]
    k = t_arr.size
    ns = np.maximum(4096, np.round(np.exp(np.log(4096.0) + u[m:m + k] *
                                          (np.log(1048576.0) - np.log(4096.0))))).astype(int)
    S = gb.Server(max_active=64, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    t0 = time.monotonic()
    nxt, done = 0, []
    while len(done) < k:
        now = time.monotonic() - t0
        while nxt < k and t_arr[nxt] <= now:
            S.submit_sdd(nxt, int(ns[nxt]), 1000003 + nxt, 3, 27, TOL)
            nxt += 1
        a, q, _ = S.status()
        if a == 0 and q == 0:
            time.sleep(0.001)
            continue
        done += S.progress(4)
        assert time.monotonic() - t0 < 120, "stream did not drain"
    assert len(done) == k and sorted(r["id"] for r in done) == list(range(k))
    assert all(r["status"] == 1 and r["rel_residual"] <= TOL for r in done)
    # requests that ran alongside others give the same bits alone
    for rid in (0, k // 3, 2 * k // 3, k - 1):
        x = S.take_result(rid, int(ns[rid]))
        S1 = gb.Server(max_active=1, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
        S1.submit_sdd(0, int(ns[rid]), 1000003 + rid, 3, 27, TOL)
        while not S1.progress(8):
            pass
        assert np.array_equal(S1.take_result(0, int(ns[rid])), x), rid
        S1.free()
    S.free()
