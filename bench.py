"""Benchmark of the sparse-solve path on one MI355X.

Default workload (the large FP64 configuration): 3D 7-point Dirichlet
Laplacian on a 160^3 grid (n = 4,096,000, nnz = 28,518,400, FP64 values
6/-1, int32 indices), x and b uniform in [-1, 1) from seed 1.  One step =
100 SpMVs y = A x, then one CG solve from x0 = 0 to relative residual 1e-10
-- the whole configuration, inputs resident in HBM before the timed region.

Prints one JSON line: value = steps (= CG solves) per second, the SpMV
roofline (algorithmic bytes 12 nnz + 4 (n+1) + 16 n per SpMV over the SpMV
kernel's average duration, from HIP events on the library stream), a short
served stream's p50/p99 latency (--served-seconds, 10 by default) and the
CPU baseline (the oracle, the C restatement of the reference's OpenMP path,
on this host's physical cores: the median of 5 runs of the 100 SpMVs and of
the full solve, and the SpMV on one thread).

--workload selects the other configurations, each printing its own line:
  zipf     irregular rows: n = 2,000,000 random symmetric diagonally
           dominant SPD matrix, degrees Zipf(1.5) clipped to [2, 4096] before
           symmetrising, seed 2; FP32-inner CG with FP64 reliable updates
           (--precision jacobi: on the Jacobi-scaled system, the default;
           mixed: unscaled) to 1e-10 against plain FP64 CG, errors against
           FP64 CG run to 1e-12
  batched  512 independent 2D/3D Laplacians (alternating), n log-uniform in
           [1000, 200000], b from seed 3 + i; all advanced together by the
           batched engine (graph-replayed steps), mixed precision to 1e-10;
           default plans against autotuned plans, and against one-at-a-time
           FP64 GPU CG
  served   Poisson arrivals (200/s for --duration s) of random SDD requests,
           n log-uniform in [4096, 1048576], degrees uniform in [3, 27]
           before symmetrising, seed per request; continuous batching with
           memory admission, mixed precision to 1e-10; p50/p99 latency, and
           the capacity on a backlog of the same requests against the plain
           one-at-a-time FP64 GPU CG
  spgemm   C = A A for the 3D 7-point Laplacian on a --spgemm-grid^3 grid
           (default 150: n = 3,375,000, 164 M products, 83 M entries of C),
           resident A in, resident C out (CSR, columns in first-appearance
           order, bitwise the reference's sums); the oracle's sequential
           product on the same matrix as the CPU baseline
  ingest   the 160^3 Laplacian as .csr text (written as the reference
           writes it, %g) read into a resident matrix: text parsed on the
           GPU; also the parse alone with the text already in HBM, and the
           binary CSR file; the oracle's reader (the reference's loop with
           the C library's strtod) as the CPU baseline
  rcm      Cuthill-McKee ordering (the CLI's -b) of the 100^3 Laplacian with
           its rows and columns scrambled by a random permutation (seed 11):
           orderings per second, the bandwidth before and after, and the
           FP64 SpMV on the scrambled and on the reordered matrix; the
           oracle's sequential ordering as the CPU baseline
  dropin   gooseberry_spmult with host pointers (upload, plan, multiply and
           copy back on every call) at the 256^2 and 160^3 Laplacians
  pagerank `gooseberry pagerank` on the Zipf graph of the zipf workload
           (n = 2,000,000, weights |a_ij|, no sinks): damping 0.85, a fixed
           100 iterations from the resident matrix (column sums,
           normalisation and the power iteration all timed) with the
           autotuned SpMV plan, and with the default plan; the oracle's
           iterations on the same graph as the CPU baseline and check
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
# --precision choices of the mixed-precision workloads
PRECISIONS = {"fp64": lambda gb: gb.PRECISION_FP64,
              "mixed": lambda gb: gb.PRECISION_MIXED,
              "jacobi": lambda gb: gb.PRECISION_MIXED_JACOBI}


def spmv_bytes(n, nnz, val_bytes=8, vec_bytes=8):
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * vec_bytes * n


def format_bytes(n, nnz, info):
    """Compulsory bytes of one SpMV in the plan's stored format, for plans
    whose rows are all LDS-staged row blocks: each entry's column (a 16-bit
    offset or a 32-bit index) and value (an 8-bit index into the value table
    or the FP64 value), the row pointers, one 16-byte descriptor per block,
    x once and y once."""
    col = 2 if info["index16"] else 4
    val = 1 if info["values8"] else 8
    return nnz * (col + val) + 4 * (n + 1) + 16 * info["nblocks_stream"] + 16 * n


def spmv_kernel_name(info):
    """The kernel the plan runs for a plain FP64 SpMV."""
    i16 = "true" if info["index16"] else "false"
    vd = "true" if info["values8"] else "false"
    if info["values8"] and not info["nrows_wave"] and not info["nrows_block"]:
        return "gb_spmv_pipe<double, false, %s, %s, 512u, 8>" % (i16, vd)
    return "gb_spmv_kernel<double, false, %s, %s>" % (i16, vd)


def plan_kernel_name(plan):
    """The FP64 SpMV kernel(s) a tuned plan runs on a matrix with long rows."""
    if plan.get("xtile"):
        return "gb_spmv_xtile (x staged through LDS in tiles of %d columns)" % plan["xtile"]
    if plan.get("csplit"):
        return ("gb_spmv_kernel<double, false, false, false> over %d column slices + "
                "gb_split_sum" % plan["csplit"])
    return "gb_spmv_kernel<double, false, false, false>"


def cg_iter_bytes(n, nnz):
    # minimum per iteration: A once, read/write x, r, p
    return 12 * nnz + 4 * (n + 1) + 48 * n


# the sources the SpMV kernel is built from: a counter summary counts only
# for the build it was measured on
SPMV_SOURCES = ["gooseberry_amd/csrc/hip/gb_spmv.hip", "gooseberry_amd/csrc/hip/gb_matrix.h",
                "gooseberry_amd/csrc/hip/gb_common.h", "gooseberry_amd/csrc/hip/gb_xtile.h",
                "gooseberry_amd/build.py"]


def spmv_source_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in SPMV_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def load_traffic():
    """Measured HBM bytes and traced duration per SpMV launch from the
    committed rocprofv3 summary, or None for each when the summary was taken
    on a different build of the kernel (its source hash differs)."""
    p = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if not os.path.exists(p):
        return None, None, "no counter summary committed"
    with open(p) as f:
        t = json.load(f)
    if t.get("sources_sha256") != spmv_source_hash():
        return None, None, "the committed counter summary is of another build of the kernel"
    return (t.get("traffic_bytes_per_launch"), t.get("trace_avg_launch_ms"),
            "profiles/spmv_traffic.json (" + t.get("source", "") + ")")


def host_cpu():
    """The host's CPU: model, sockets, physical cores and logical CPUs in
    this process's affinity mask, and the cgroup CPU quota (if any)."""
    info = {"model": None, "sockets": None, "physical_cores": None,
            "logical_cpus": len(os.sched_getaffinity(0)), "cgroup_cpus": None}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    info["model"] = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    mask = os.sched_getaffinity(0)
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True,
                             text=True, timeout=20).stdout
        cores, socks = set(), set()
        for ln in out.splitlines():
            if ln.startswith("#") or not ln.strip():
                continue
            cpu, core, sock = (int(v) if v else 0 for v in ln.split(",")[:3])
            if cpu in mask:
                cores.add((sock, core))
                socks.add(sock)
        info["physical_cores"] = len(cores) or None
        info["sockets"] = len(socks) or None
    except Exception:
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                info["cgroup_cpus"] = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    return info


def cpu_threads(host):
    """All physical cores this process may use: the affinity mask's physical
    cores, capped by the cgroup quota and by OMP_NUM_THREADS when the
    environment sets it (the machine's share of the host)."""
    t = host["physical_cores"] or host["logical_cpus"]
    for cap in (host["cgroup_cpus"], int(os.environ.get("OMP_NUM_THREADS", "0")) or None):
        if cap:
            t = min(t, cap)
    return max(1, t)


def cpu_baseline(rp, ri, rv, x, b, tol_abs, reps=5, nspmv=NSPMV, single_spmv=True,
                 what="the 160^3 Laplacian"):
    """The oracle (the C + OpenMP restatement of the reference's path, -O3,
    no FMA) on this host's cores, over the reference's "Compute" regions
    (src/gooseberry_bin.c:1697-1802 for `blas -r 100`, :2078-2091 for `cgd`):
    nspmv back-to-back SpMVs, and one full CG solve to the absolute
    tolerance tol_abs (work vectors allocated inside, each iteration's line
    printed to /dev/null); the median of `reps` runs of each, on all
    physical cores (OMP_PLACES=cores, OMP_PROC_BIND=close), plus the SpMV on
    one thread.  A step = nspmv SpMVs + one solve, as on the GPU."""
    # the OpenMP runtime reads these when the oracle is first loaded
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    host = host_cpu()
    threads = cpu_threads(host)
    n, nnz = rp.size - 1, int(rp[-1])

    def spmv_runs(k):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(k):
                orc.spmv(rp, ri, rv, x)
            ts.append((time.perf_counter() - t0) / k)
        return float(np.median(ts))

    orc.set_threads(threads)
    orc.spmv(rp, ri, rv, x)  # first touch, page faults
    t_spmv = spmv_runs(nspmv) if nspmv else spmv_runs(20)
    solves, iters, rel = [], None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        xs, it, rmse, _ = orc.cgd(rp, ri, rv, b, tol_abs, print_devnull=True)
        solves.append(time.perf_counter() - t0)
        iters, rel = int(it), float(rmse / np.linalg.norm(b))
    t_solve = float(np.median(solves))
    out = {
        "value": 1.0 / (nspmv * t_spmv + t_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": ("the oracle (C + OpenMP restatement of src/blas.c and src/cgd.c, -O3, no "
                   "FMA) on %s: median of %d runs of %d back-to-back SpMVs and of %d full CG "
                   "solves (work vectors allocated inside, per-iteration line to /dev/null) on "
                   "%d threads (OMP_PLACES=%s, OMP_PROC_BIND=%s); a step = %d SpMVs + 1 solve"
                   % (what, reps, nspmv or 20, reps, threads, os.environ.get("OMP_PLACES"),
                      os.environ.get("OMP_PROC_BIND"), nspmv)),
        "host": host,
        "spmv_gbps": spmv_bytes(n, nnz) / t_spmv / 1e9,
        "spmv_ms": t_spmv * 1e3,
        "solve_ms": t_solve * 1e3,
        "solve_ms_runs": [v * 1e3 for v in solves],
        "cg_iterations": iters,
        "cg_final_rel_residual": rel,
        "cg_iter_ms": t_solve * 1e3 / max(1, iters),
    }
    if single_spmv:
        orc.set_threads(1)
        t1 = spmv_runs(nspmv or 20)
        out["single_thread_spmv_ms"] = t1 * 1e3
        out["single_thread_spmv_gbps"] = spmv_bytes(n, nnz) / t1 / 1e9
        orc.set_threads(threads)
    return out


def run_laplacian(args, gb):
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
    traffic, trace_ms, traffic_src = load_traffic()
    minfo = M.info()
    fb = format_bytes(n, nnz, minfo)

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
            "kernel": spmv_kernel_name(minfo),
            "algorithmic_bytes_per_launch": ab,
            "avg_launch_ms": avg_spmv_ms,
            "timing": "HIP events around 100 back-to-back launches on the library stream",
            "trace_avg_launch_ms": trace_ms,
            "trace_frac": (ab / (trace_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS) if trace_ms else None,
            "traffic_source": traffic_src,
            # the plan reads each entry as a 16-bit column offset and an 8-bit
            # index into the value table (3 bytes instead of the CSR's 12), so
            # the CSR bytes over the launch time can exceed the HBM peak; the
            # fractions of the bytes the format holds and of the measured
            # traffic are the ones the hardware bounds
            "format_bytes_per_launch": fb,
            "format_frac": fb / (avg_spmv_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS,
            "traffic_frac": (traffic / (avg_spmv_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS
                             if traffic else None),
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
    del M, dx, dy, db, dsol
    if args.served_seconds > 0:
        sv = served_leg(args, gb, args.rate, args.served_seconds)
        out["served_p50_latency_ms"] = sv["p50_latency_ms"]
        out["served_p99_latency_ms"] = sv["p99_latency_ms"]
        out["served"] = sv
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rp, ri, rv, x, b, TOL * float(np.linalg.norm(b)),
                                           reps=args.cpu_reps)
    else:
        out["cpu_baseline"] = None
    return out


def served_leg(args, gb, rate, duration):
    """A short served stream beside the default workload: Poisson arrivals
    at `rate`/s for `duration` s of the served workload's requests through
    the continuous-batching server (64 slots, --precision), latency from
    arrival to result."""
    t_arr, ns, seeds = served_requests(gb, rate, duration)
    S = gb.Server(max_active=64, precision=PRECISIONS[args.precision](gb))
    done, lat, wall = serve(gb, S, t_arr, ns, seeds)
    S.free()
    lats = np.array([lat[r["id"]] for r in done]) * 1e3
    return {"workload": ("Poisson arrivals %.0f/s for %.0f s (%d requests): random SDD SPD, n "
                         "log-uniform in [4096, 1048576], degrees U[3,27] then symmetrised, "
                         "generated on the GPU at admission; 64 slots, precision %s, to "
                         "relative residual %g" % (rate, duration, len(ns), args.precision, TOL)),
            "requests": int(len(ns)), "solves_per_s": len(done) / wall,
            "p50_latency_ms": pct(lats, 50), "p99_latency_ms": pct(lats, 99),
            "all_converged": all(r["status"] == 1 and r["rel_residual"] <= TOL for r in done),
            "max_rel_residual": float(max(r["rel_residual"] for r in done)) if done else None}


# ------------------------------------------------------------- shared bits

def line(value, unit, args, ms_per_step, dtype, config, **extra):
    out = {
        "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic", "config": config,
    }
    out.update(extra)
    return out


def true_rel_residual(gb, M, db, dx, bh=None):
    """||b - A x|| / ||b|| with the FP64 SpMV on the device."""
    dy = gb.DeviceArray(M.nrows)
    M.spmv_device(dx, dy)
    b = db.numpy() if bh is None else bh
    r = b - dy.numpy()
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def spmv_roofline(gb, M, dx, dy, nrep=20, val_bytes=8):
    """Average FP64 SpMV launch time from HIP events on the library stream."""
    e0, e1 = gb.Event(), gb.Event()
    for _ in range(3):
        M.spmv_device(dx, dy)
    e0.record()
    for _ in range(nrep):
        M.spmv_device(dx, dy)
    e1.record()
    ms = e0.elapsed_ms(e1) / nrep
    inf = M.info()
    ab = spmv_bytes(inf["nrows"], int(inf["nnz"]), val_bytes)
    a = ab / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": a, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": a / PEAK_HBM_GBPS, "traffic": None,
            "kernel": plan_kernel_name(inf["plan"]),
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": ms}


def pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else None


# ------------------------------------------------ irregular rows (Zipf)

def run_zipf(args, gb):
    n = args.zipf_n
    t0 = time.perf_counter()
    M = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    gen_s = time.perf_counter() - t0
    inf = M.info()
    nnz = int(inf["nnz"])
    db = gb.uniform_device(2, n, stream=1)
    bh = db.numpy()
    log("zipf: n=%d nnz=%d (mean row %.2f) generated in %.2f s" % (n, nnz, nnz / n, gen_s))
    dxs = gb.uniform_device(2, n, stream=0)
    dys = gb.DeviceArray(n)
    roof_default = spmv_roofline(gb, M, dxs, dys)
    plan, tuned_ms = M.autotune(gb.PRECISION_FP64, 5)
    roof = spmv_roofline(gb, M, dxs, dys)
    log("zipf: SpMV default plan %.1f us, autotuned %.1f us (%s)"
        % (roof_default["avg_launch_ms"] * 1e3, roof["avg_launch_ms"] * 1e3, plan))
    del dxs, dys

    # the FP64 reference solution: FP64 CG run to 1e-12
    dref = gb.DeviceArray(n)
    iref = M.cgd_device(db, dref, 1e-12, gb.TOL_RELATIVE)
    xref = dref.numpy()
    log("zipf: FP64 CG to 1e-12: %d iterations, status %d" % (iref.niter, iref.status))

    def errors_of(dx, info):
        x = dx.numpy()
        return {"status": int(info.status), "iterations": int(info.niter),
                "refreshes": int(info.nrefine),
                "true_rel_residual": true_rel_residual(gb, M, db, dx, bh),
                "rel_error_vs_fp64_1e-12": float(np.linalg.norm(x - xref) /
                                                 np.linalg.norm(xref))}

    def fp64_solve():
        dx = gb.DeviceArray(n)
        gb.synchronize()
        t = time.perf_counter()
        info = M.cgd_device(db, dx, TOL, gb.TOL_RELATIVE)
        return time.perf_counter() - t, info, dx

    prec = PRECISIONS[args.precision](gb)

    def mixed_solve(precision=prec):
        dx = gb.DeviceArray(n)
        B = gb.Batch(1)
        gb.synchronize()
        t = time.perf_counter()
        slot = B.add(M, db, dx, TOL, gb.TOL_RELATIVE, 0, precision)
        B.solve()
        dt = time.perf_counter() - t
        info = B.info(slot)
        B.free()
        return dt, info, dx

    # each solve on the plan tuned for its own SpMV: FP64 CG on the FP64
    # plan, the FP32-inner solves on the plan tuned for the FP32 SpMV (the
    # batched engine's kernels)
    for _ in range(args.warmup):
        fp64_solve()
    f64 = [fp64_solve() for _ in range(args.steps)]
    f64_errs = errors_of(f64[-1][2], f64[-1][1])
    plan32, tuned32_ms = M.autotune(gb.PRECISION_MIXED, 5)
    log("zipf: FP32-inner plan %s (FP32 SpMV %.1f us)" % (plan32, tuned32_ms * 1e3))
    for _ in range(args.warmup):
        mixed_solve()
    mix = [mixed_solve() for _ in range(args.steps)]
    log("zipf: FP64 %.1f ms, mixed %.1f ms" % (np.median([a for a, _, _ in f64]) * 1e3,
                                               np.median([a for a, _, _ in mix]) * 1e3))


    t64 = float(np.median([a for a, _, _ in f64]))
    tmx = float(np.median([a for a, _, _ in mix]))
    other = {}
    for name in ("mixed", "jacobi"):
        p = PRECISIONS[name](gb)
        if p != prec:
            mixed_solve(p)
            dt, info, dx = mixed_solve(p)
            other[name] = dict(errors_of(dx, info), solve_ms=dt * 1e3)
    cfg = {"workload": ("random symmetric diagonally dominant SPD, n=%d, degrees Zipf(1.5) "
                        "clipped to [2,4096] then symmetrised (nnz=%d, mean row length "
                        "%.2f), off-diagonals U[-1,0), diagonal 1.01 x |row sum|, seed 2; "
                        "b U[-1,1) seed 2; step = one FP32-inner / FP64-refined CG solve "
                        "(precision %s) to relative residual %g"
                        % (n, nnz, nnz / n, args.precision, TOL)),
           "n": n, "nnz": nnz, "tol": TOL, "precision": args.precision}
    out = line(args.steps / sum(a for a, _, _ in mix), "solves/s", args, tmx * 1e3,
               "f32+f64", cfg, roofline=roof,
               mixed=dict(errors_of(mix[-1][2], mix[-1][1]), solve_ms=tmx * 1e3),
               fp64=dict(f64_errs, solve_ms=t64 * 1e3),
               speedup_mixed_vs_fp64=t64 / tmx, other_precisions=other,
               p50_solve_ms=pct([a * 1e3 for a, _, _ in mix], 50),
               p99_solve_ms=pct([a * 1e3 for a, _, _ in mix], 99),
               fp64_reference={"tol": 1e-12, "iterations": int(iref.niter)},
               autotune={"plan": plan, "spmv_ms_default": roof_default["avg_launch_ms"],
                         "spmv_ms_tuned": roof["avg_launch_ms"],
                         "fp32_inner_plan": plan32, "fp32_spmv_ms_tuned": tuned32_ms},
               generate_s=gen_s)
    if not args.no_cpu_baseline:
        rp, ri, rv = M.download()
        x = gb.uniform(2, n, stream=0)
        cb = cpu_baseline(rp, ri, rv, x, bh, TOL * float(np.linalg.norm(bh)), reps=3,
                          nspmv=0, single_spmv=False,
                          what="the Zipf matrix (n=%d, nnz=%d)" % (n, nnz))
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    return out


def run_zipf_spmv(args, gb):
    """The Zipf matrix's FP64 SpMV alone: the default plan, then the autotuned
    plan (each 20 launches timed with HIP events) -- the short run the
    counter profiles of the irregular gather use."""
    n = args.zipf_n
    M = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    nnz = int(M.info()["nnz"])
    dx = gb.uniform_device(2, n, stream=0)
    dy = gb.DeviceArray(n)
    roof_default = spmv_roofline(gb, M, dx, dy)
    if args.csplit is None and args.xtile is None:
        plan, _ = M.autotune(gb.PRECISION_FP64, 5)
    else:  # a fixed plan (counter runs: the profiler's slowdown skews tuning)
        M.replan({"csplit": args.csplit} if args.csplit else {"xtile": args.xtile})
        plan = M.info()["plan"]
    roof = spmv_roofline(gb, M, dx, dy)
    roof["kernel"] = plan_kernel_name(plan)
    cfg = {"workload": ("FP64 SpMV of the random SDD SPD matrix, n=%d, degrees Zipf(1.5) clipped "
                        "to [2,4096] then symmetrised (nnz=%d), x U[-1,1) seed 2" % (n, nnz)),
           "n": n, "nnz": nnz}
    return line(roof["achieved"], "GB/s", args, roof["avg_launch_ms"], "f64", cfg,
                roofline=roof, default_plan=roof_default, autotuned_plan=plan,
                cpu_baseline=None)


# ------------------------------------------------------ batched systems

def batched_systems(gb, nsys):
    """512 (by default) 2D/3D Laplacians, alternating, n log-uniform in
    [1000, 200000] from the library's host generator (seed 3, stream 2)."""
    u = gb.uniform(3, nsys, 0.0, 1.0, stream=2)
    out = []
    for i in range(nsys):
        dims = 2 if i % 2 == 0 else 3
        target = float(np.exp(np.log(1000.0) + u[i] * (np.log(200000.0) - np.log(1000.0))))
        side = max(2, int(round(target ** (1.0 / dims))))
        out.append((dims, side))
    return out


def fp64_references(gb, mats, rhs):
    """Every system's FP64 reference solution: the plain FP64 CG (the
    reference's loop on the GPU) run to relative residual 1e-12."""
    refs = []
    for M, db in zip(mats, rhs):
        dx = gb.DeviceArray(M.nrows)
        M.cgd_device(db, dx, 1e-12, gb.TOL_RELATIVE)
        refs.append(dx.numpy())
        dx.free()
    return refs


def result_errors(gb, mats, rhs, xs, refs):
    """Every result's true relative residual (FP64 SpMV on the device) and
    relative error against its FP64 reference solution."""
    res, err = [], []
    for M, db, dx, xr in zip(mats, rhs, xs, refs):
        res.append(true_rel_residual(gb, M, db, dx))
        err.append(float(np.linalg.norm(dx.numpy() - xr) / np.linalg.norm(xr)))
    return {"systems": len(res), "true_rel_residual_max": max(res),
            "true_rel_residual_median": float(np.median(res)),
            "rel_error_vs_fp64_1e-12_max": max(err),
            "rel_error_vs_fp64_1e-12_median": float(np.median(err))}


def run_batched(args, gb):
    shapes = batched_systems(gb, args.nsys)
    mats, rhs = [], []
    t0 = time.perf_counter()
    for i, (dims, side) in enumerate(shapes):
        rp, ri, rv = gb.laplacian(dims, side)
        M = gb.Matrix(rp, ri, rv)
        mats.append(M)
        rhs.append(gb.DeviceArray.from_host(gb.uniform(3 + i, M.nrows, stream=1)))
    ns = np.array([M.nrows for M in mats])
    nnz = int(sum(int(M.info()["nnz"]) for M in mats))
    log("batched: %d systems, sum n=%d, nnz=%d, built in %.1f s"
        % (len(mats), ns.sum(), nnz, time.perf_counter() - t0))
    xs = [gb.DeviceArray(M.nrows) for M in mats]
    prec = PRECISIONS[args.precision](gb)

    def batch_solve(precision):
        B = gb.Batch(len(mats))
        gb.synchronize()
        t = time.perf_counter()
        slots = [B.add(M, b, x, TOL, gb.TOL_RELATIVE, 0, precision)
                 for M, b, x in zip(mats, rhs, xs)]
        B.solve()
        dt = time.perf_counter() - t
        infos = [B.info(s) for s in slots]
        B.free()
        return dt, infos

    def summary(dt, infos):
        its = [int(i.niter) for i in infos]
        return {"solves_per_s": len(infos) / dt, "wall_ms": dt * 1e3,
                "all_converged": all(int(i.status) == 1 for i in infos),
                "iterations_max": max(its), "iterations_mean": float(np.mean(its)),
                "refreshes_mean": float(np.mean([int(i.nrefine) for i in infos]))}

    nnzs = np.array([int(M.info()["nnz"]) for M in mats], dtype=np.float64)

    def inner_bytes(infos):
        """FP32 inner iterations' algorithmic bytes, each system's own count
        (8 nnz + 4 (n+1) + 24 n per iteration; the FP64 refreshes not
        counted, so a lower bound)"""
        its = np.array([int(i.niter) for i in infos], dtype=np.float64)
        return float(np.sum(its * (8 * nnzs + 4 * (ns + 1) + 24 * ns)))

    # the FP64 references (CG to 1e-12) and the plain FP64 GPU CG one solve
    # at a time over every system (timed)
    t = time.perf_counter()
    refs = fp64_references(gb, mats, rhs)
    log("batched: FP64 references (CG to 1e-12) in %.1f s" % (time.perf_counter() - t))
    single_iters = []
    if not args.no_single:
        gb.synchronize()
        t1 = time.perf_counter()
        for M, b, x in zip(mats, rhs, xs):
            single_iters.append(int(M.cgd_device(b, x, TOL, gb.TOL_RELATIVE).niter))
        gb.synchronize()
        t_single = time.perf_counter() - t1
    for _ in range(args.warmup):
        batch_solve(prec)
    runs = [batch_solve(prec) for _ in range(args.steps)]
    errs = result_errors(gb, mats, rhs, xs, refs)
    other = {}
    for name in ("fp64", "mixed", "jacobi"):
        p = PRECISIONS[name](gb)
        if p != prec:
            other[name] = summary(*batch_solve(p))
    # autotuned plans: one tuning per distinct matrix shape class
    t = time.perf_counter()
    tuned = gb.autotune_all(mats, prec)
    tune_s = time.perf_counter() - t
    runs_tuned = [batch_solve(prec) for _ in range(args.steps)]
    for M in mats:
        M.replan(None)

    tm = float(np.median([a for a, _ in runs]))
    ab = inner_bytes(runs[-1][1])
    achieved = ab / runs[-1][0] / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
            "kernel": "whole batched step (gb_b_spmv, gb_b_reduce, gb_b_update, gb_b_pupdate "
                      "over all systems): FP32 inner-iteration bytes of every system over the "
                      "solve's wall time",
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": runs[-1][0] * 1e3}
    cfg = {"workload": ("%d independent Laplacians (2D 5-point / 3D 7-point alternating), n "
                        "log-uniform in [1000, 200000] (sum n=%d, nnz=%d), b U[-1,1) seed "
                        "3+i; step = all systems solved together by the batched engine, "
                        "precision %s, to relative residual %g"
                        % (len(mats), ns.sum(), nnz, args.precision, TOL)),
           "systems": len(mats), "sum_n": int(ns.sum()), "nnz": nnz, "tol": TOL,
           "precision": args.precision}
    out = line(len(mats) * args.steps / sum(a for a, _ in runs), "solves/s", args,
               tm * 1e3, "f32+f64" if args.precision != "fp64" else "f64", cfg,
               batched=summary(*runs[-1]), every_result=errs, other_precisions=other,
               roofline=roof, cpu_baseline=None)
    if single_iters:
        out["one_at_a_time_fp64"] = {"solves_per_s": len(mats) / t_single,
                                     "systems": len(mats), "wall_ms": t_single * 1e3,
                                     "iterations_mean": float(np.mean(single_iters))}
        out["speedup_vs_one_at_a_time_fp64"] = (len(mats) / tm) / (len(mats) / t_single)
    out["autotuned_plans"] = dict(summary(*runs_tuned[-1]), seconds=tune_s,
                                  classes=tuned)
    return out


# --------------------------------------------------------- served stream

def served_requests(gb, rate, duration):
    """Poisson arrivals at `rate`/s for `duration` s (seed 5): arrival times,
    n log-uniform in [4096, 1048576], request seeds 1000003 + k."""
    m = int(rate * duration * 1.5) + 64
    u = gb.uniform(5, 2 * m, 0.0, 1.0, stream=3)
    gaps = -np.log1p(-u[:m]) / rate
    t = np.cumsum(gaps)
    t = t[t < duration]
    k = t.size
    ns = np.exp(np.log(4096.0) + u[m:m + k] * (np.log(1048576.0) - np.log(4096.0)))
    ns = np.maximum(4096, np.round(ns)).astype(np.int64)
    return t, ns, 1000003 + np.arange(k)


def now_mono():
    return time.clock_gettime(time.CLOCK_MONOTONIC)


def serve(gb, S, t_arr, ns, seeds, steps=4, progress_every=10.0):
    """Submit each request at its arrival time, advance the server, return
    per-request latency (done - arrival) and the done records."""
    k = len(t_arr)
    t0 = now_mono()
    nxt, done, lat = 0, [], {}
    last_log = t0
    while len(done) < k:
        now = now_mono() - t0
        while nxt < k and t_arr[nxt] <= now:
            S.submit_sdd(int(nxt), int(ns[nxt]), int(seeds[nxt]), 3, 27, TOL)
            nxt += 1
        a, q, _ = S.status()
        if a == 0 and q == 0:
            if nxt < k:
                time.sleep(max(0.0, min(0.002, t_arr[nxt] - (now_mono() - t0))))
            continue
        for r in S.progress(steps):
            done.append(r)
            lat[r["id"]] = r["done_s"] - (t0 + t_arr[r["id"]])
        if now_mono() - last_log > progress_every:
            last_log = now_mono()
            log("served: t=%.1f s submitted %d done %d active %d queued %d"
                % (last_log - t0, nxt, len(done), a, q))
    return done, lat, now_mono() - t0


def run_served(args, gb):
    t_arr, ns, seeds = served_requests(gb, args.rate, args.duration)
    log("served: %d requests over %.0f s, mean n %.0f" % (len(ns), args.duration, ns.mean()))
    # the Poisson stream through the continuous-batching server
    prec = PRECISIONS[args.precision](gb)
    S = gb.Server(max_active=64, precision=prec)
    done, lat, wall = serve(gb, S, t_arr, ns, seeds)
    S.free()
    lats = np.array([lat[r["id"]] for r in done]) * 1e3
    ok = all(r["status"] == 1 and r["rel_residual"] <= TOL for r in done)
    nnz = np.array([r["nnz"] for r in done])
    its = np.array([r["niter"] for r in done])

    # capacity on a backlog of the first requests, inputs resident in HBM
    # before the timed regions: the plain FP64 CG one solve at a time, then
    # the same systems through the server (mixed precision, 64 slots)
    kc = min(args.capacity_requests, len(ns))
    if kc == 0:
        return line(len(done) / wall, "solves/s", args, wall * 1e3, "f32+f64",
                    {"workload": "served stream (no capacity comparison)",
                     "requests": int(len(ns))},
                    p50_latency_ms=pct(lats, 50), p99_latency_ms=pct(lats, 99),
                    all_converged=ok, roofline=None, cpu_baseline=None)
    mats = [gb.generate_sdd(int(ns[i]), int(seeds[i]), 3, 27) for i in range(kc)]
    bs = [gb.uniform_device(int(seeds[i]), int(ns[i]), stream=1) for i in range(kc)]
    xs = [gb.DeviceArray(int(ns[i])) for i in range(kc)]
    gb.synchronize()
    t = time.perf_counter()
    single_its = []
    for i in range(kc):
        info = mats[i].cgd_device(bs[i], xs[i], TOL, gb.TOL_RELATIVE)
        single_its.append(int(info.niter))
    gb.synchronize()
    t_single = time.perf_counter() - t
    refs = fp64_references(gb, mats, bs)
    xs2 = [gb.DeviceArray(int(ns[i])) for i in range(kc)]
    S = gb.Server(max_active=64, precision=prec)
    gb.synchronize()
    t = time.perf_counter()
    for i in range(kc):
        S.submit(i, mats[i], bs[i], xs2[i], TOL)
    cap_done = []
    while len(cap_done) < kc:
        cap_done += S.progress(4)
    t_served = time.perf_counter() - t
    S.free()
    cap_ok = all(r["status"] == 1 and r["rel_residual"] <= TOL for r in cap_done)
    cap_errs = result_errors(gb, mats, bs, xs2, refs)
    del mats, bs, xs, xs2, refs
    cap_served = kc / t_served
    cap_single = kc / t_single
    # whole-run roofline of the served capacity: every request's FP32
    # inner-iteration bytes (8 nnz + 4 (n+1) + 24 n per iteration, the FP64
    # refreshes not counted) over the run's wall time
    ab = float(sum(int(r["niter"]) * (8 * int(r["nnz"]) + 4 * (int(r["n"]) + 1)
                                       + 24 * int(r["n"])) for r in cap_done))
    achieved = ab / t_served / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
            "kernel": "whole served capacity run (batched SpMV, reductions and updates over "
                      "all admitted requests): FP32 inner-iteration bytes over its wall time",
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": t_served * 1e3}
    log("served: capacity %.1f solves/s batched mixed vs %.1f one-at-a-time FP64"
        % (cap_served, cap_single))

    cfg = {"workload": ("served stream: Poisson arrivals %.0f/s for %.0f s (%d requests), "
                        "random SDD SPD n log-uniform in [4096, 1048576], degrees U[3,27] "
                        "then symmetrised, values as the Zipf workload, seed per request; "
                        "continuous batching (64 slots, memory admission), FP32-inner / "
                        "FP64-refined CG (precision %s) to relative residual %g"
                        % (args.rate, args.duration, len(ns), args.precision, TOL)),
           "requests": int(len(ns)), "rate": args.rate, "duration_s": args.duration,
           "tol": TOL}
    return line(len(done) / wall, "solves/s", args, wall * 1e3, "f32+f64", cfg,
                p50_latency_ms=pct(lats, 50), p99_latency_ms=pct(lats, 99),
                all_converged=ok, mean_nnz=float(nnz.mean()), mean_iterations=float(its.mean()),
                max_rel_residual=float(max(r["rel_residual"] for r in done)),
                capacity={"requests": kc, "inputs": "resident in HBM before timing",
                          "served_mixed_solves_per_s": cap_served,
                          "one_at_a_time_fp64_solves_per_s": cap_single,
                          "ratio": cap_served / cap_single,
                          "served_all_converged": cap_ok,
                          "every_result": cap_errs,
                          "served_iterations_mean": float(np.mean([r["niter"] for r in cap_done])),
                          "one_at_a_time_mean_iterations": float(np.mean(single_its))},
                roofline=roof, cpu_baseline=None)


# ------------------------------------------------ sparse x sparse (SpGEMM)

def spgemm_counts(rp, ri):
    """Number of products of A A: each entry (i, l) of A meets row l."""
    lens = np.diff(rp.astype(np.int64))
    return int(lens[ri.astype(np.int64)].sum())


def run_spgemm(args, gb):
    g = args.spgemm_grid
    rp, ri, rv = gb.laplacian(3, g)
    n, nnz = rp.size - 1, int(rp[-1])
    products = spgemm_counts(rp, ri)
    A = gb.Matrix(rp, ri, rv)
    e0, e1 = gb.Event(), gb.Event()
    nnz_c = None
    for _ in range(args.warmup):
        Cm = A.spmultsp(A)
        nnz_c = Cm.nnz
        Cm.free()
    gb.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0.record()
        Cm = A.spmultsp(A)
        e1.record()
        dev_ms.append(e0.elapsed_ms(e1))
        nnz_c = Cm.nnz
        Cm.free()
    gb.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    ab = 2 * (12 * nnz + 4 * (n + 1)) + 12 * nnz_c + 4 * (n + 1)
    achieved = ab / (ms * 1e-3) / 1e9
    extra = {
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
            "kernel": "whole product (expansion, two stable radix sorts, run sums, "
                      "first-appearance sort, output, plan of C)",
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": float(np.mean(dev_ms)),
        },
        "products": products, "nnz_c": nnz_c,
        "products_per_s": products / (ms * 1e-3),
    }
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.set_threads(1)
        t = time.perf_counter()
        k = 0
        while True:
            orc.spmultsp_sp((rp, ri, rv), (rp, ri, rv), n)
            k += 1
            if time.perf_counter() - t > args.cpu_seconds or k >= 5:
                break
        t = (time.perf_counter() - t) / k
        cpu = {"value": 1.0 / t, "unit": "spgemm/s", "cores": 1, "kind": "port",
               "sample": "%d sequential products A A of the same %d^3 Laplacian by the "
                         "oracle (C restatement of src/blas.c:124-234, count then fill, "
                         "-O3, no FMA) on one core" % (k, g),
               "products_per_s": products / t}
    extra["cpu_baseline"] = cpu
    return line(args.steps / elapsed * 1.0, "spgemm/s", args, ms, "f64",
                {"workload": "C = A A, 3D 7-point Laplacian %d^3 (n=%d, nnz=%d), %d products, "
                             "C nnz %d; resident A in, resident C out" % (g, n, nnz, products,
                                                                          nnz_c),
                 "n": n, "nnz": nnz}, **extra)


# ------------------------------------------------ CSR ingest

def run_ingest(args, gb):
    import ctypes as C
    import tempfile
    g = args.grid
    rp, ri, rv = gb.laplacian(3, g)
    n, nnz = rp.size - 1, int(rp[-1])
    d = tempfile.mkdtemp(prefix="gb_ingest_")
    csr, bcsr = os.path.join(d, "lap.csr"), os.path.join(d, "lap.bcsr")
    try:
        gb.write_csr(csr, rp, ri, rv, n)
        gb.write_csr(bcsr, rp, ri, rv, n, fmt=gb.FORMAT_BCSR)
        text_bytes = os.path.getsize(csr)
        for _ in range(args.warmup):
            gb.Matrix.read(csr).free()
        gb.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            M = gb.Matrix.read(csr)
            M.free()
        gb.synchronize()
        elapsed = time.perf_counter() - t0
        ms = elapsed / args.steps * 1e3
        # the binary file
        t = time.perf_counter()
        for _ in range(args.steps):
            gb.Matrix.read(bcsr, gb.FORMAT_BCSR).free()
        bin_ms = (time.perf_counter() - t) / args.steps * 1e3
        # the parse alone, text resident in HBM
        data = np.fromfile(csr, dtype=np.uint8)
        dt = gb.DeviceArray(data.size, np.uint8)
        dt.upload(data)
        e0, e1 = gb.Event(), gb.Event()
        parse_ms = []
        for rep in range(args.steps + 1):
            nr, nc = C.c_uint32(), C.c_uint32()
            p = [C.c_void_p(), C.c_void_p(), C.c_void_p()]
            e0.record()
            gb.check(gb.lib().gooseberry_parse_csr_device(
                dt.ptr, data.size, C.byref(nr), C.byref(nc), C.byref(p[0]), C.byref(p[1]),
                C.byref(p[2])), "gooseberry_parse_csr_device")
            e1.record()
            if rep:
                parse_ms.append(e0.elapsed_ms(e1))
            for q in p:
                gb.lib().gooseberry_device_free(q.value)
        dt.free()
        pms = float(np.mean(parse_ms))
        ab = text_bytes + 4 * (n + 1) + 12 * nnz   # text in, CSR out
        achieved = ab / (pms * 1e-3) / 1e9
        # the library's own host reader and the oracle's, same file
        t = time.perf_counter()
        gb.read_sparse(csr)
        host_s = time.perf_counter() - t
        cpu = None
        if not args.no_cpu_baseline:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as orc
            k, t = 0, time.perf_counter()
            while True:
                orc.read_sparse(csr)
                k += 1
                if time.perf_counter() - t > args.cpu_seconds or k >= 3:
                    break
            t = (time.perf_counter() - t) / k
            cpu = {"value": text_bytes / t / 1e6, "unit": "MB/s of text", "cores": 1,
                   "kind": "port",
                   "sample": "%d reads of the same %.0f MB file by the oracle's reader (C "
                             "restatement of src/io.c:369-498: getline, strtoull, strtod) "
                             "on one core" % (k, text_bytes / 1e6)}
    finally:
        for f in (csr, bcsr):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)
    return line(text_bytes / (ms * 1e-3) / 1e6, "MB/s of text", args, ms, "u8",
                {"workload": ".csr text of the 3D 7-point Laplacian %d^3 (n=%d, nnz=%d, %.0f "
                             "MB, %%g values) read into a resident matrix; file in the page "
                             "cache" % (g, n, nnz, text_bytes / 1e6),
                 "n": n, "nnz": nnz, "text_bytes": text_bytes},
                roofline={"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS,
                          "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                          "kernel": "whole parse with the text in HBM (line split, row "
                                    "flags, count and fill passes, scans)",
                          "algorithmic_bytes_per_launch": ab, "avg_launch_ms": pms},
                parse_in_hbm_gbps=text_bytes / (pms * 1e-3) / 1e9,
                binary_read_ms=bin_ms, binary_bytes=24 + 4 * (n + 1) + 12 * nnz,
                host_reader_ms=host_s * 1e3, cpu_baseline=cpu)


# ------------------------------------------------ Cuthill-McKee

def bandwidth(rp, ri):
    rows = np.repeat(np.arange(rp.size - 1, dtype=np.int64), np.diff(rp.astype(np.int64)))
    return int(np.max(np.abs(rows - ri.astype(np.int64)))) if ri.size else 0


def run_rcm(args, gb):
    g = 100
    rp, ri, rv = gb.laplacian(3, g)
    n, nnz = rp.size - 1, int(rp[-1])
    q = np.random.default_rng(11).permutation(n).astype(np.uint32)
    srp, sri, srv = gb.permute_sparse(rp, ri, rv, n, q, q)
    M = gb.Matrix(srp, sri, srv)
    for _ in range(args.warmup):
        M.cuthillmckee()
    gb.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        perm = M.cuthillmckee()
    gb.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    prp, pri, prv = gb.permute_sparse(srp, sri, srv, n, perm, perm)
    M2 = gb.Matrix(prp, pri, prv)
    dx, dy = gb.DeviceArray.from_host(gb.uniform(1, n)), gb.DeviceArray(n)
    before = spmv_roofline(gb, M, dx, dy)
    after = spmv_roofline(gb, M2, dx, dy)
    ab = 4 * (n + 1) + 4 * nnz + 4 * n          # structure read once, perm written
    achieved = ab / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        k, t = 0, time.perf_counter()
        while True:
            ref = orc.cuthillmckee(srp, sri)
            k += 1
            if time.perf_counter() - t > args.cpu_seconds or k >= 3:
                break
        t = (time.perf_counter() - t) / k
        assert np.array_equal(ref, perm)
        cpu = {"value": 1.0 / t, "unit": "orderings/s", "cores": 1, "kind": "port",
               "sample": "%d orderings of the same scrambled matrix by the oracle (C "
                         "restatement of src/permute.c:49-110 with a binary heap) on one "
                         "core; its permutation equals the GPU's" % k}
    return line(args.steps / elapsed, "orderings/s", args, ms, "u32",
                {"workload": "Cuthill-McKee of the 3D 7-point Laplacian %d^3 (n=%d, nnz=%d) "
                             "with rows and columns scrambled (seed 11)" % (g, n, nnz),
                 "n": n, "nnz": nnz},
                roofline={"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS,
                          "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                          "kernel": "whole ordering (degree, sort, claim steps, radix sorts, "
                                    "one-workgroup stepping)",
                          "algorithmic_bytes_per_launch": ab, "avg_launch_ms": ms},
                bandwidth_before=bandwidth(srp, sri), bandwidth_after=bandwidth(prp, pri),
                spmv_scrambled_ms=before["avg_launch_ms"], spmv_reordered_ms=after["avg_launch_ms"],
                spmv_scrambled_gbps=before["achieved"], spmv_reordered_gbps=after["achieved"],
                cpu_baseline=cpu)


# ------------------------------------------------------------ PageRank

PR_ITERS = 100


def run_pagerank(args, gb):
    n = args.zipf_n
    M0 = gb.generate_spd(n, 2, gb.DEGREE_ZIPF, 1.5, 2, 4096)
    rp, ri, rv = M0.download()
    del M0
    rv = np.abs(rv)
    nnz = int(rp[-1])
    M = gb.Matrix(rp, ri, rv)
    log("pagerank: graph n=%d nnz=%d (mean row %.2f)" % (n, nnz, nnz / n))
    dout = gb.DeviceArray(n)

    def timed(steps, warmup):
        for _ in range(warmup):
            M.pagerank_device(dout, 0.85, PR_ITERS, 0.0)
        gb.synchronize()
        t0 = time.perf_counter()
        done = 0
        for _ in range(steps):
            it, rmse, nsinks = M.pagerank_device(dout, 0.85, PR_ITERS, 0.0)
            done += it
        gb.synchronize()
        elapsed = time.perf_counter() - t0
        ms = elapsed / steps * 1e3
        # with minerror 0 a run still stops early when an iterate repeats its
        # predecessor exactly (the error is checked every 10 iterations), so
        # the iterations counted are those the runs made.  The per-iteration
        # cost: the same run capped at one iteration removes the preparation
        # (column sums, normalisation, sinks)
        M.pagerank_device(dout, 0.85, 1, 0.0)
        gb.synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            M.pagerank_device(dout, 0.85, 1, 0.0)
        gb.synchronize()
        ms1 = (time.perf_counter() - t1) / 3 * 1e3
        per_run = done / steps
        return elapsed, ms, (ms - ms1) / max(per_run - 1, 1), ms1, nsinks, done

    _, ms_default, it_ms_default, _, _, it_default = timed(1, min(1, args.warmup))
    plan, _ = M.autotune(gb.PRECISION_FP64, 5)
    log("pagerank: default plan %.3f ms per iteration; autotuned plan %s"
        % (it_ms_default, plan))
    elapsed, ms, it_ms, ms1, nsinks, done = timed(args.steps, args.warmup)
    # one iteration: the SpMV over the normalised values (12 nnz + 4 (n+1)
    # + 16 n) and out = out d + c (16 n)
    ab = 12 * nnz + 4 * (n + 1) + 32 * n
    achieved = ab / (it_ms * 1e-3) / 1e9
    cpu = None
    check = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        orc.set_threads(threads)
        nval, sinks = orc.pagerank_prepare(rp, ri, rv, n)
        t = time.perf_counter()
        orc.pagerank(rp, ri, nval, sinks, 0.85, 2, 0.0)
        t = (time.perf_counter() - t) / 2
        k = int(max(2, min(50, args.cpu_seconds * 0.7 / max(t, 1e-6))))
        t = time.perf_counter()
        ref, _, _ = orc.pagerank(rp, ri, nval, sinks, 0.85, k, 0.0)
        t = time.perf_counter() - t
        got = gb.DeviceArray(n)
        M.pagerank_device(got, 0.85, k, 0.0)
        g = got.numpy()
        check = float(np.abs(g - ref).max() / np.abs(ref).max())
        assert check <= 1e-12, check
        cpu = {"value": k / t, "unit": "iterations/s", "cores": threads, "kind": "port",
               "sample": "%d power iterations of the oracle (C + OpenMP restatement of "
                         "src/gooseberry_bin.c:2622-2737) on %d threads on the same graph, "
                         "column sums and normalisation excluded; its vector matches the "
                         "GPU's to %.1e relative" % (k, threads, check)}
    return line(done / elapsed, "iterations/s", args, ms, "f64",
                {"workload": "PageRank (damping 0.85, at most %d iterations, minimum error 0) "
                             "of the Zipf(1.5) graph "
                             "n=%d nnz=%d (random symmetric pattern, degrees clipped to "
                             "[2,4096] before symmetrising, seed 2), weights |a_ij|; "
                             "autotuned SpMV plan" % (PR_ITERS, n, nnz),
                 "n": n, "nnz": nnz, "iterations": PR_ITERS},
                iterations_run=done / args.steps, default_plan_iterations_run=it_default,
                edges_per_s=nnz * done / elapsed,
                iteration_ms=it_ms, preparation_ms=ms1 - it_ms, sinks=nsinks,
                plan=plan, default_plan_ms_per_step=ms_default,
                default_plan_iteration_ms=it_ms_default,
                roofline={"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS,
                          "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                          "kernel": "one power iteration (SpMV over the normalised values, "
                                    "sink sum, update)",
                          "algorithmic_bytes_per_launch": ab, "avg_launch_ms": it_ms},
                oracle_check_rel=check, cpu_baseline=cpu)


# --------------------------------------- the host-pointer drop-in boundary

def run_dropin(args, gb):
    """gooseberry_spmult with host pointers (src/gooseberry.c:53-67): each call
    uploads A and x, builds A's SpMV plan, multiplies and copies y back.  Per
    call time at the CPU-runnable 256^2 Laplacian and the 160^3 Laplacian,
    beside the resident SpMV and the oracle's SpMV on this host.  Every GPU
    call is timed before the oracle's OpenMP runtime starts: under
    OMP_PROC_BIND it binds this thread to one core, and a host-pointer call
    made from a thread bound to one core is a different measurement."""
    cases = []
    for dims, side in ((2, 256), (3, args.grid)):
        rp, ri, rv = gb.laplacian(dims, side)
        n = rp.size - 1
        x = gb.uniform(1 if dims == 3 else 0, n, stream=0 if dims == 3 else 1)
        y = np.empty(n)
        for _ in range(max(1, args.warmup)):
            gb.spmult(n, n, 1, rp, ri, rv, x, crowval=y)
        ts = []
        for _ in range(args.steps):
            t = time.perf_counter()
            gb.spmult(n, n, 1, rp, ri, rv, x, crowval=y)
            ts.append(time.perf_counter() - t)
        M = gb.Matrix(rp, ri, rv)
        dx, dy = gb.DeviceArray.from_host(x), gb.DeviceArray(n)
        roof = spmv_roofline(gb, M, dx, dy)
        del M, dx, dy
        cases.append((dims, side, rp, ri, rv, x, y, ts, roof))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle as orc
    host = host_cpu()
    orc.set_threads(cpu_threads(host))
    res = {}
    for dims, side, rp, ri, rv, x, y, ts, roof in cases:
        n = rp.size - 1
        assert np.array_equal(y, orc.spmv(rp, ri, rv, x)) or dims == 3
        tc = []
        for _ in range(5):
            t = time.perf_counter()
            orc.spmv(rp, ri, rv, x)
            tc.append(time.perf_counter() - t)
        ab = spmv_bytes(n, int(rp[-1]))
        res["%dd_%d" % (dims, side)] = {
            "n": n, "nnz": int(rp[-1]), "host_bytes_in": int(4 * (n + 1) + 12 * int(rp[-1]) + 8 * n),
            "dropin_ms_per_call": float(np.median(ts)) * 1e3,
            "resident_spmv_ms": roof["avg_launch_ms"],
            "oracle_spmv_ms": float(np.median(tc)) * 1e3,
            "dropin_algorithmic_gbps": ab / float(np.median(ts)) / 1e9}
    big = res["3d_%d" % args.grid]
    return line(1e3 / big["dropin_ms_per_call"], "calls/s", args, big["dropin_ms_per_call"],
                "f64", {"workload": "gooseberry_spmult with host pointers (A, x uploaded, plan "
                        "built, y copied back on every call): 2D 256^2 and 3D %d^3 Laplacians"
                        % args.grid}, per_config=res, roofline=None,
                cpu_baseline={"value": 1e3 / big["oracle_spmv_ms"], "unit": "calls/s",
                              "cores": cpu_threads(host), "kind": "port", "host": host,
                              "sample": "5 oracle SpMVs of the 3D Laplacian (median)"})


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 10; zipf 3; batched 3)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps (default 2; zipf 1; batched 1)")
    ap.add_argument("--workload", default="laplacian",
                    choices=["laplacian", "zipf", "batched", "served", "spgemm", "ingest",
                             "rcm", "pagerank", "dropin", "zipf-spmv"])
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--zipf-n", type=int, default=2000000)
    ap.add_argument("--nsys", type=int, default=512)
    ap.add_argument("--csplit", type=int, default=None,
                    help="zipf-spmv: this many column slices instead of the autotuned plan")
    ap.add_argument("--xtile", type=int, default=None,
                    help="zipf-spmv: x tiles of this many columns instead of the autotuned plan")
    ap.add_argument("--spgemm-grid", type=int, default=150)
    ap.add_argument("--duration", type=float, default=60.0,
                    help="served: seconds of Poisson arrivals")
    ap.add_argument("--rate", type=float, default=200.0,
                    help="served: arrivals per second")
    ap.add_argument("--capacity-requests", type=int, default=200,
                    help="served: requests of the backlog capacity comparison")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="jacobi",
                    help="precision of the zipf, batched and served solves: plain mixed "
                         "(FP32 inner, FP64 refreshes) or mixed on the Jacobi-scaled system")
    ap.add_argument("--served-seconds", type=float, default=10.0,
                    help="length of the served stream beside the default workload (0: none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="batched: skip the one-at-a-time FP64 comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU time bound of the other workloads' baselines")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="runs of each CPU baseline leg (the median is reported)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 4 device copies of exactly 256 MiB, the "
                         "known-size kernel that scales the rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build is single-GPU (the path "
              "would shard by row blocks with a halo exchange). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    defaults = {"laplacian": (10, 2), "zipf": (3, 1), "batched": (3, 1), "served": (1, 0),
                "spgemm": (5, 1), "ingest": (5, 1), "rcm": (3, 1), "pagerank": (3, 1),
                "dropin": (5, 1), "zipf-spmv": (1, 0)}
    st, wu = defaults[args.workload]
    args.steps = st if args.steps is None else args.steps
    args.warmup = wu if args.warmup is None else args.warmup

    import gooseberry_amd as gb

    run = {"laplacian": run_laplacian, "zipf": run_zipf, "batched": run_batched,
           "served": run_served, "spgemm": run_spgemm, "ingest": run_ingest,
           "rcm": run_rcm, "pagerank": run_pagerank, "dropin": run_dropin,
           "zipf-spmv": run_zipf_spmv}[args.workload]
    out = run(args, gb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
