"""Benchmark of the sparse-solve path on one MI355X.

Workload (the large FP64 configuration): 3D 7-point Dirichlet Laplacian on a
160^3 grid (n = 4,096,000, nnz = 28,518,400, FP64 values 6/-1, int32
indices), x and b uniform in [-1, 1) from seed 1.  One step = 100 SpMVs
y = A x, then one CG solve from x0 = 0 to relative residual 1e-10 -- the
whole configuration, inputs resident in HBM before the timed region.

Prints one JSON line: value = steps (= CG solves) per second, the SpMV
roofline (algorithmic bytes 12 nnz + 4 (n+1) + 16 n per SpMV over the SpMV
kernel's average duration, from HIP events on the library stream) and the
CPU baseline (the oracle, the C restatement of the reference's OpenMP path,
timed on a bounded sample on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CG solves/sec and SpMV HBM GB/s (% of peak); p50/p99 solve latency"
PEAK_HBM_GBPS = 8000.0
GRID = 160
TOL = 1e-10
NSPMV = 100


def spmv_bytes(n, nnz, val_bytes=8, vec_bytes=8):
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * vec_bytes * n


def cg_iter_bytes(n, nnz):
    # minimum per iteration: A once, read/write x, r, p
    return 12 * nnz + 4 * (n + 1) + 48 * n


def load_traffic():
    """Measured HBM bytes per SpMV launch from the committed rocprofv3
    counter summary (None when it has not been measured)."""
    p = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def cpu_baseline(rp, ri, rv, x, b, gpu_iters, seconds=15.0):
    """The oracle on this host's cores: a bounded sample of SpMVs and CG
    iterations; solves/s extrapolated with the full solve's iteration
    count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    orc.set_threads(threads)
    n, nnz = rp.size - 1, int(rp[-1])
    # SpMV sample
    orc.spmv(rp, ri, rv, x)
    t0 = time.perf_counter()
    k = 0
    while True:
        orc.spmv(rp, ri, rv, x)
        k += 1
        if time.perf_counter() - t0 > seconds / 3 or k >= NSPMV:
            break
    t_spmv = (time.perf_counter() - t0) / k
    # CG sample: a fixed number of iterations of the reference's loop
    iters = 5
    t0 = time.perf_counter()
    orc.cgd(rp, ri, rv, b, 0.0, niter=iters)
    t_it = (time.perf_counter() - t0) / iters
    budget = seconds - (time.perf_counter() - t0) - t_spmv * k
    if budget > 4 * t_it * iters:
        more = max(iters, min(200, int(budget / t_it)))
        t0 = time.perf_counter()
        orc.cgd(rp, ri, rv, b, 0.0, niter=more)
        t_it = (time.perf_counter() - t0) / more
        iters = more
    t_solve = NSPMV * t_spmv + gpu_iters * t_it
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": ("%d SpMVs and %d CG iterations of the 160^3 Laplacian with the oracle "
                   "(C + OpenMP restatement of src/blas.c and src/cgd.c, -O3, no FMA) on "
                   "%d threads; a step's time = 100 SpMVs + %d iterations (the full "
                   "solve's count) at the sampled per-iteration time"
                   % (k, iters, threads, gpu_iters)),
        "spmv_gbps": spmv_bytes(n, nnz) / t_spmv / 1e9,
        "spmv_ms": t_spmv * 1e3,
        "cg_iter_ms": t_it * 1e3,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 4 device copies of exactly 256 MiB, the "
                         "known-size kernel that scales the rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build is single-GPU (the path "
              "would shard by row blocks with a halo exchange; see DESIGN.md). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gooseberry_amd as gb

    g = args.grid
    rp, ri, rv = gb.laplacian(3, g)
    n, nnz = rp.size - 1, int(rp[-1])
    x = gb.uniform(1, n, stream=0)
    b = gb.uniform(1, n, stream=1)

    M = gb.Matrix(rp, ri, rv)
    dx = gb.DeviceArray.from_host(x)
    dy = gb.DeviceArray(n)
    db = gb.DeviceArray.from_host(b)
    dsol = gb.DeviceArray(n)
    e0, e1 = gb.Event(), gb.Event()

    def step():
        e0.record()
        for _ in range(NSPMV):
            M.spmv_device(dx, dy)
        e1.record()
        spmv_ms = e0.elapsed_ms(e1) / NSPMV  # waits for the SpMVs
        t0 = time.perf_counter()
        info = M.cgd_device(db, dsol, TOL, gb.TOL_RELATIVE)
        t1 = time.perf_counter()  # cgd_device returns after its final poll
        return spmv_ms, t1 - t0, info

    if args.calibrate_copy:
        src = gb.DeviceArray(32 << 20)  # 256 MiB of float64
        dst = gb.DeviceArray(32 << 20)
        for _ in range(4):
            dst.copy_from(src)
        gb.synchronize()
        del src, dst

    for _ in range(args.warmup):
        step()
    gb.synchronize()
    spmv_ms, solve_s, iters, status = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, s, info = step()
        spmv_ms.append(a)
        solve_s.append(s)
        iters.append(int(info.niter))
        status.append(int(info.status))
    gb.synchronize()
    elapsed = time.perf_counter() - t0

    avg_spmv_ms = float(np.mean(spmv_ms))
    ab = spmv_bytes(n, nnz)
    achieved = ab / (avg_spmv_ms * 1e-3) / 1e9
    it = int(np.median(iters))
    solve_ms = np.array(solve_s) * 1e3
    cg_iter_ms = float(np.median(solve_ms)) / max(it, 1)
    traffic = load_traffic()

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("3D 7-point Laplacian %d^3 (n=%d, nnz=%d), FP64 values, int32 "
                         "indices, x/b U[-1,1) seed 1; step = %d SpMVs then CG from x0=0 to "
                         "relative residual %g" % (g, n, nnz, NSPMV, TOL)),
            "n": n, "nnz": nnz, "tol": TOL, "spmv_per_step": NSPMV,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": traffic,
            "kernel": "gb_spmv_kernel<double,false>",
            "algorithmic_bytes_per_launch": ab,
            "avg_launch_ms": avg_spmv_ms,
        },
        "spmv_gbps": achieved,
        "spmv_frac_of_peak": achieved / PEAK_HBM_GBPS,
        "cg_iterations": it,
        "cg_status": status,
        "p50_solve_ms": float(np.percentile(solve_ms, 50)),
        "p99_solve_ms": float(np.percentile(solve_ms, 99)),
        "cg_iter_ms": cg_iter_ms,
        "cg_iter_min_bytes_gbps": cg_iter_bytes(n, nnz) / (cg_iter_ms * 1e-3) / 1e9,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rp, ri, rv, x, b, it, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
