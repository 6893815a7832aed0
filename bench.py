"""Benchmark: FP64 BCSR SpMV block-size sweep on one MI355X.

Workload (the SpMV sweep the headline metric is quoted on): 9-point block
stencil matrices (diagonally dominant, seed 0) for bs in {1,2,3,4,6,8,12,16,
24,32}, square grid side n = round(sqrt(S/(9*bs^2))) with S = 4.4e8 stored
values ("hbm" sizing, ~3.5-6.3 GB per matrix, all resident in HBM together) or
S = 4.4e7 ("literal" sizing), x uniform[0,1].

One step = one FP64 SpMV on every matrix of the sweep, each with its fastest
kernel family/layout (chosen by timing every family first).  value = compulsory
bytes of the step / step time (GB/s); the timed region is K steps between two
HIP events on the work stream, inputs already resident in HBM.

Also reported: per-(bs, layout, family) GB/s (with the CSR expansion of each
matrix under the three CSR families), the dominant kernel's roofline
(HIP-event launch time; compulsory bytes per launch; HBM traffic per launch
read from the counter profile of that kernel's symbol under profiles/, null
when the profile has no row for it), the solver and serving legs, and the
CPU baseline: the oracle's single-threaded CSR SpMV (main.cpp:36-45
restated) on every sweep matrix, CSR / BCSR SpMV and BiCGStab on the 256x256
bs=2 system, and the FP64 BiCGStab on the 2048x1024 bs=4 system, run in a
child process pinned to one core that never opens the GPU.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

BS_SWEEP = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)
CPU_CHILD_FLAG = "--cpu-child"


def grid_side(bs, total_values):
    return int(round(math.sqrt(total_values / (9.0 * bs * bs))))


# ---- CPU baseline: a child process on one pinned core, never opening the GPU --
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_child(spec):
    """The oracle's CPU path (test infrastructure, timed here only as the
    baseline), single-threaded, pinned to spec["core"]:
    - 256x256 bs=2 stencil: CSR SpMV (main.cpp:36-45 order), BCSR SpMV, mean
      of 10 runs each (x uniform[0,1], the GPU's x), and the FP64 BiCGStab of
      b = A*1 to 1e-8 (bicgstab.cu:388-513 restated), one full solve;
    - 2048x1024 bs=4 convective system, b = A x*: the FP64 BiCGStab to 1e-10,
      one full solve; its solution goes to spec["x_path"];
    - every sweep matrix: CSR SpMV of its CSR expansion, mean of 3-10 runs
      (as many as spec["seconds"] / #bs allows)."""
    os.sched_setaffinity(0, {spec["core"]})
    import oracle
    L = oracle.lib()

    def clog(msg):
        print(f"[bench cpu] {msg}", file=sys.stderr, flush=True)

    def mean_time(fn, reps):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    def csr_bytes(c):
        return 12.0 * c.nnz + 4.0 * (c.n_rows + 1) + 16.0 * c.n_rows

    out = {"cpu_model": cpu_model(), "core": spec["core"]}
    # -- 256x256 bs=2 --------------------------------------------------------
    m = oracle.gen_stencil9(256, 256, 2, 0, 0)
    c = oracle.bcsr_to_csr(m)
    x = oracle.rng_unit_vec(0, 4, c.n_cols)
    y = np.empty(c.n_rows)
    L.orc_csr_spmv(c.n_rows, c.row_ptr, c.columns, c.values, x, y)
    t_csr = mean_time(lambda: L.orc_csr_spmv(c.n_rows, c.row_ptr, c.columns, c.values, x, y), 10)
    L.orc_bcsr_spmv(m.n_rows, m.bs, 0, m.row_ptr, m.columns, m.values, x, y)
    t_bcsr = mean_time(lambda: L.orc_bcsr_spmv(m.n_rows, m.bs, 0, m.row_ptr, m.columns, m.values, x, y), 10)
    bcsr_bytes = 8.0 * m.nnzb * m.bs ** 2 + 4.0 * m.nnzb + 4.0 * (m.n_rows + 1) + 16.0 * c.n_rows
    b = oracle.csr_spmv(c, np.ones(c.n_rows))
    t0 = time.perf_counter()
    xo, info = oracle.bicgstab(c, b, 1e-8, 10000)
    t_solve = time.perf_counter() - t0
    out["system_256x256_bs2"] = {
        "csr_spmv_us": round(t_csr * 1e6, 2), "csr_GBs": round(csr_bytes(c) / t_csr / 1e9, 2),
        "csr_GFLOPs": round(2.0 * c.nnz / t_csr / 1e9, 3),
        "bcsr_spmv_us": round(t_bcsr * 1e6, 2), "bcsr_GBs": round(bcsr_bytes / t_bcsr / 1e9, 2),
        "bicgstab_ms": round(t_solve * 1e3, 2), "iterations": info.iterations,
        "max_abs_x_minus_1": float(np.max(np.abs(xo - 1.0)))}
    clog(f"256x256 bs=2: CSR {t_csr*1e6:.0f} us, BCSR {t_bcsr*1e6:.0f} us, BiCGStab {t_solve*1e3:.0f} ms "
         f"({info.iterations} it)")
    del m, c, x, y, b, xo
    # -- 2048x1024 bs=4 FP64 BiCGStab ------------------------------------------
    m = oracle.gen_stencil9(2048, 1024, 4, 1, 0)
    c = oracle.bcsr_to_csr(m)
    del m
    xs = oracle.rng_sym_vec(1, 5, c.n_rows)
    b = oracle.csr_spmv(c, xs)
    t0 = time.perf_counter()
    xo, info = oracle.bicgstab(c, b, 1e-10, spec["max_iter"])
    t_solve = time.perf_counter() - t0
    np.save(spec["x_path"], xo)
    out["system_2048x1024_bs4"] = {
        "bicgstab_seconds": round(t_solve, 3), "iterations": info.iterations, "status": info.status,
        "rel_residual": info.rel_residual,
        "err_vs_xstar": float(np.linalg.norm(xo - xs) / np.linalg.norm(xs))}
    clog(f"2048x1024 bs=4 FP64 BiCGStab: {info.iterations} it in {t_solve:.2f} s")
    del c, xs, b, xo
    # -- the sweep matrices, CSR ---------------------------------------------
    per_bs = {}
    budget = spec["seconds"] / max(len(spec["bs"]), 1)
    for bs in spec["bs"]:
        side = grid_side(bs, spec["total_values"])
        m = oracle.gen_stencil9(side, side, bs, 0, 0)
        c = oracle.bcsr_to_csr(m)
        del m
        x = oracle.rng_unit_vec(0, 4, c.n_cols)
        y = np.empty(c.n_rows)
        t1 = mean_time(lambda: L.orc_csr_spmv(c.n_rows, c.row_ptr, c.columns, c.values, x, y), 1)
        reps = int(min(10, max(3, budget // max(t1, 1e-9))))
        t = mean_time(lambda: L.orc_csr_spmv(c.n_rows, c.row_ptr, c.columns, c.values, x, y), reps)
        per_bs[str(bs)] = {"us": round(t * 1e6, 1), "reps": reps, "csr_GBs": round(csr_bytes(c) / t / 1e9, 2),
                           "GFLOPs": round(2.0 * c.nnz / t / 1e9, 3), "nnz": c.nnz}
        clog(f"sweep bs={bs}: CSR SpMV {t*1e3:.1f} ms ({per_bs[str(bs)]['csr_GBs']} GB/s, mean of {reps})")
        del c, x, y
    out["sweep_csr"] = per_bs
    return out


if __name__ == "__main__" and len(sys.argv) > 2 and sys.argv[1] == CPU_CHILD_FLAG:
    print(json.dumps(cpu_child(json.loads(sys.argv[2]))), flush=True)
    sys.exit(0)


def start_cpu_child(bs_list, total_values, seconds, max_iter):
    """Starts the CPU leg on the last core this process may use and keeps
    this process off that core.  Called before anything opens the GPU."""
    avail = sorted(os.sched_getaffinity(0))
    core = avail[-1]
    if len(avail) > 1:
        os.sched_setaffinity(0, set(avail[:-1]))
    spec = {"core": core, "bs": bs_list, "total_values": total_values, "seconds": seconds, "max_iter": max_iter,
            "x_path": os.path.join(tempfile.gettempdir(), f"bench_cpu_x_{os.getpid()}.npy")}
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), CPU_CHILD_FLAG, json.dumps(spec)],
                            stdout=subprocess.PIPE, text=True)
    return proc, spec, len(avail)


import block_matrix_format_performance_amd as bm  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of the SpMV kernels (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, scaled by a copy-kernel calibration), one table row per mangled
# kernel symbol and sizing
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "spmv_hbm_traffic_pmc.md")
FAMILY_NAMES = {
    bm.BCSR_THREAD_PER_ROW: "thread_per_row",
    bm.BCSR_GROUP_PER_ROW: "group_per_row",
    bm.BCSR_SHUFFLE_X: "shuffle_x",
    bm.BCSR_BLOCK_ROW_GROUP: "block_row_group",
    bm.BCSR_LIBRARY: "rocsparse_bsrmv",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(symbol, sizing):
    """HBM bytes per dispatch of `symbol` on the `sizing` sweep matrix from
    the counter profile's table (rows `| symbol | sizing | bs | ... |
    traffic bytes | ...`), or None when the profile has no such row."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            lines = f.readlines()
    except OSError:
        return None
    header = None
    for line in lines:
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if line.lstrip().startswith("| symbol"):
            header = cells
            continue
        if header and len(cells) == len(header) and cells[0] == symbol and cells[1] == sizing:
            return float(cells[header.index("traffic bytes")])
    return None


def build_case(bs, side, layouts):
    """The sweep matrix for one bs, generated in device memory in each layout."""
    return {layout: bm.Matrix.stencil(side, side, bs, 0, seed=0, dtype=bm.F64, layout=layout)
            for layout in layouts}


def measure_copy(nbytes, reps, stream):
    """Streaming copy kernel (nbytes read + nbytes written) and read-only
    kernel (nbytes read): the measured-bandwidth yardsticks."""
    nbytes = int(nbytes) // 16 * 16
    src = bm.DeviceArray(nbytes // 8)
    dst = bm.DeviceArray(nbytes // 8)
    src.fill(2, value=1.0)
    L = bm.lib()
    for _ in range(3):
        bm.check(L.bmfp_stream_copy(dst.ptr, src.ptr, nbytes, stream.handle))
    e0, e1 = bm.Event(), bm.Event()
    e0.record(stream.handle)
    for _ in range(reps):
        bm.check(L.bmfp_stream_copy(dst.ptr, src.ptr, nbytes, stream.handle))
    e1.record(stream.handle)
    t = e1.elapsed_since(e0) / reps
    for _ in range(3):
        bm.check(L.bmfp_stream_read(src.ptr, nbytes, dst.ptr, stream.handle))
    e0.record(stream.handle)
    for _ in range(reps):
        bm.check(L.bmfp_stream_read(src.ptr, nbytes, dst.ptr, stream.handle))
    e1.record(stream.handle)
    tr = e1.elapsed_since(e0) / reps
    src.free()
    dst.free()
    return {"bytes_moved": 2 * nbytes, "us": round(t * 1e6, 2), "GB/s": round(2 * nbytes / t / 1e9, 1),
            "read_only": {"bytes_read": nbytes, "us": round(tr * 1e6, 2), "GB/s": round(nbytes / tr / 1e9, 1)}}


def cpu_baseline_result(proc, spec, host_cpus, step_bytes_by_bs, gpu_value, timeout):
    """Collects the CPU leg and states it in the sweep's unit: the sweep
    step's compulsory BCSR bytes over the summed CPU CSR SpMV times (the
    same numerator as the GPU value)."""
    try:
        text, _ = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        proc.kill()
        proc.communicate()
        log("[bench] CPU leg did not finish in time")
        return None, None
    if proc.returncode != 0 or not text.strip():
        log(f"[bench] CPU leg failed (exit {proc.returncode})")
        return None, None
    res = json.loads(text.strip().splitlines()[-1])
    x_cpu = None
    try:
        x_cpu = np.load(spec["x_path"])
        os.remove(spec["x_path"])
    except OSError:
        pass
    sweep = res["sweep_csr"]
    common = [bs for bs in step_bytes_by_bs if str(bs) in sweep]
    t_cpu = sum(sweep[str(bs)]["us"] for bs in common) * 1e-6
    value = sum(step_bytes_by_bs[bs] for bs in common) / t_cpu / 1e9 if t_cpu > 0 else None
    out = {"value": round(value, 3) if value else None, "unit": "GB/s", "cores": 1, "kind": "port",
           "sample": "oracle CSR SpMV (main.cpp:36-45 order) of the CSR expansion of every sweep matrix at the "
                     f"{'hbm' if spec['total_values'] > 1e8 else 'literal'} sizing, mean of 3-10 runs each; "
                     "value = the sweep step's compulsory BCSR bytes / summed CPU times; one pinned core "
                     "(child process without the GPU)",
           "cpu_model": res["cpu_model"], "host_cpus": host_cpus,
           "gpu_over_cpu": round(gpu_value / value, 1) if value else None,
           "per_bs_csr": sweep,
           "system_256x256_bs2": res["system_256x256_bs2"],
           "system_2048x1024_bs4": res["system_2048x1024_bs4"]}
    return out, x_cpu


def run_only(args, total_values, stream):
    """--only BS:LAYOUT:FAMILY[,BS:LAYOUT:FAMILY...]: each kernel on its sweep
    matrix (built, timed --steps times, freed in turn), one JSON line each
    with the kernel's symbol; then the copy yardstick."""
    for item in args.only.split(","):
        bs_s, lay_s, fam_s = item.split(":")
        bs = int(bs_s)
        layout = bm.ROW_MAJOR if lay_s == "row" else bm.COL_MAJOR
        fam = {v: k for k, v in FAMILY_NAMES.items()}[fam_s]
        side = grid_side(bs, total_values)
        A = build_case(bs, side, (layout,))[layout]
        x = bm.DeviceArray(A.x_len)
        x.fill(0, seed=0, stream_id=4)
        y = bm.DeviceArray(A.y_len)
        nbytes, _ = A.traffic()
        t = A.spmv_timed(x, y, fam, warmup=args.warmup, reps=args.steps, stream=stream.handle)
        sym = A.kernel_symbol(fam) if fam != bm.BCSR_LIBRARY else None
        print(json.dumps({"kernel": item, "symbol": sym, "side": side, "launches": args.warmup + args.steps,
                          "us": round(t * 1e6, 2), "algorithmic_bytes": nbytes,
                          "GB/s": round(nbytes / t / 1e9, 1)}), flush=True)
        A.destroy()
        x.free()
        y.free()
    if args.copy_mb > 0:
        print(json.dumps({"copy_yardstick": measure_copy(args.copy_mb * 1e6, 5, stream)}), flush=True)


def config0_gpu(stream, reps):
    """The 256x256 bs=2 stencil on the GPU: SpMV with the AUTO family (x
    uniform[0,1]) and the FP64 BiCGStab of b = A*1 to 1e-8, x0 = 0."""
    A = bm.Matrix.stencil(256, 256, 2, 0, seed=0)
    n = A.y_len
    x = bm.DeviceArray(n)
    x.fill(0, seed=0, stream_id=4)
    y = bm.DeviceArray(n)
    nbytes, flops = A.traffic()
    t = A.spmv_timed(x, y, bm.SPMV_AUTO, warmup=5, reps=reps, stream=stream.handle)
    ones = bm.DeviceArray(n)
    ones.fill(2, value=1.0)
    b = bm.DeviceArray(n)
    A.spmv(ones, b)
    S = bm.BiCGStab(A)
    S.solve(b, x, 1e-8, 2)                           # builds the iteration graph
    info = S.solve(b, x, 1e-8, 10000)
    err = float(np.max(np.abs(x.download() - 1.0)))
    S.destroy()
    A.destroy()
    return {"spmv_us": round(t * 1e6, 2), "spmv_GBs": round(nbytes / t / 1e9, 1),
            "spmv_GFLOPs": round(flops / t / 1e9, 2), "bicgstab_ms": round(info.seconds * 1e3, 3),
            "iterations": info.iterations, "max_abs_x_minus_1": err}


def solver_comparison(stream, max_iter):
    """FP64 BiCGStab vs FP64-refinement around FP32 / FP16+scale inner BiCGStab
    on the 2048x1024 bs=4 convective block system (off-diagonal blocks
    -(K+0.3C), diagonal 1.02*rowsum+1), b = A x*, x* uniform[-1,1] (seed 1),
    x0 = 0, to 1e-10 relative residual."""
    nx, ny, bs = 2048, 1024, 4
    t0 = time.perf_counter()
    n = nx * ny
    A = bm.Matrix.stencil(nx, ny, bs, 1, seed=0)
    xs = bm.DeviceArray(n * bs)
    xs.fill(1, seed=1, stream_id=5)
    b = bm.DeviceArray(n * bs)
    A.spmv(xs, b)
    bm.synchronize()
    h_xs = xs.download()
    log(f"[bench] solver system {nx}x{ny} bs={bs} n={n * bs} built in {time.perf_counter() - t0:.1f} s")
    res = {"system": f"{nx}x{ny} nodes, bs={bs}, convective 9-point stencil, n={n * bs}, "
                     f"nnzb={A.nnzb}, eps=1e-10, x0=0, b=A x*"}

    def err(x, ref):
        return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))

    x64 = bm.DeviceArray(n * bs)
    S = bm.BiCGStab(A)
    S.solve(b, x64, 1e-10, 2)                       # builds the iteration graph
    info = S.solve(b, x64, 1e-10, max_iter)
    h64 = x64.download()
    res["fp64"] = {"seconds": round(info.seconds, 5), "iterations": info.iterations,
                   "status": info.status, "rel_residual": info.true_rel_residual,
                   "err_vs_xstar": err(h64, h_xs)}
    log(f"[bench] FP64 BiCGStab: {info.iterations} it, {info.seconds*1e3:.1f} ms, "
        f"res {info.true_rel_residual:.2e}")
    S.destroy()
    sols = {"fp64": h64}
    for name, inner, tol in (("mixed_fp32", bm.F32, 1e-5), ("mixed_fp16", bm.F16, 1e-3)):
        # the one-time FP64 -> FP32 / FP16+scale copy the refinement handle
        # makes of A: its device time (the copy kernels, allocation excluded)
        R = bm.Refine(A, inner)
        t_copy = R.copy_seconds()
        x = bm.DeviceArray(n * bs)
        R.solve(b, x, 1e-10, tol, 1, 2)             # builds the inner graphs
        ri = R.solve(b, x, 1e-10, tol, 50, max_iter)
        hx = x.download()
        res[name] = {"seconds": round(ri.seconds, 5), "outer_iterations": ri.outer_iterations,
                     "inner_iterations": ri.inner_iterations, "status": ri.status,
                     "escalations": ri.escalations, "rel_residual": ri.true_rel_residual,
                     "err_vs_fp64_solution": err(hx, h64), "err_vs_xstar": err(hx, h_xs),
                     "time_vs_fp64": round(ri.seconds / info.seconds, 4),
                     "copy_seconds": round(t_copy, 5),
                     "time_vs_fp64_with_copy": round((ri.seconds + t_copy) / info.seconds, 4),
                     "inner_dtype": {bm.F32: "f32", bm.F16: "f16"}.get(ri.inner_dtype, ri.inner_dtype)}
        sols[name] = hx
        log(f"[bench] {name}: {ri.outer_iterations} outer / {ri.inner_iterations} inner, "
            f"{ri.seconds*1e3:.1f} ms ({ri.seconds / info.seconds:.3f}x FP64; with the {t_copy*1e3:.2f} ms copy "
            f"{(ri.seconds + t_copy) / info.seconds:.3f}x), res {ri.true_rel_residual:.2e}")
        R.destroy()
    A.destroy()
    return res, sols


SERVE_STREAM = 20000
SERVE_BS = (2, 3, 4, 6, 8)


def serving_stream(seed=2):
    """The request stream: 20,000 requests from `seed`.  Per request: square
    grid side U{16..512}, bs from {2,3,4,6,8}, 1-4 right-hand sides
    uniform[-1,1], tolerance 1e-10; unit-rate exponential inter-arrival gaps
    (divided by the arrival rate)."""
    rng = np.random.default_rng(seed)
    side = rng.integers(16, 513, size=SERVE_STREAM)
    bs = rng.choice(SERVE_BS, size=SERVE_STREAM)
    nrhs = rng.integers(1, 5, size=SERVE_STREAM)
    arrivals = np.cumsum(rng.exponential(1.0, size=SERVE_STREAM))
    reqs = [{"side": int(side[i]), "bs": int(bs[i]), "nrhs": int(nrhs[i]),
             "seed": (seed << 32) + i, "rhs_seed": (seed << 32) + i} for i in range(SERVE_STREAM)]
    return reqs, arrivals - arrivals[0]


def plain_one_at_a_time(reqs, max_iter, stream, use_jacobi=False):
    """Each request served alone by the plain FP64 BiCGStab (x0 = 0, 1e-10),
    its right-hand sides one after another, with point-Jacobi preconditioning
    when use_jacobi (as gpu_bicgstab(A, true)).  Timed per request, as the
    served loop is charged: the generation of its FP64 matrix on the device
    (bmfp_stencil_generate, HIP events) and the device time of each solve
    (info.seconds: first kernel to the solution copy, host round trips
    between iteration batches included).  Not timed: allocations, solver
    set-up and graph capture, and the FP64 true-residual check each solve
    runs after it has stopped.  Compulsory bytes per iteration: two FP64
    SpMVs (matrix, x read, y written) and 15 further FP64 vector passes (two
    more with the inverse diagonal)."""
    total, per_req, xs, iters, moved, gen_s = 0.0, [], [], [], 0.0, 0.0
    e0, e1 = bm.Event(), bm.Event()
    for r in reqs:
        A = bm.Matrix.stencil(r["side"], r["side"], r["bs"], 0, seed=r["seed"])
        n = A.y_len
        B = bm.DeviceArray(r["nrhs"] * n)
        B.fill(1, seed=r["rhs_seed"], stream_id=6)
        x = bm.DeviceArray(n)
        S = bm.BiCGStab(A, use_jacobi=use_jacobi)
        S.solve(B.view(0, n), x, 1e-10, 2)            # graph capture, untimed
        bm.synchronize()
        e0.record(stream.handle)
        A.regenerate(0, r["seed"], stream.handle)     # the request's matrix, as admission builds it
        e1.record(stream.handle)
        t_req = e1.elapsed_since(e0)
        gen_s += t_req
        xr = []
        for k in range(r["nrhs"]):
            info = S.solve(B.view(k * n, n), x, 1e-10, max_iter, stream=stream.handle)
            if info.status != bm.OK:
                raise RuntimeError(f"plain FP64 BiCGStab failed on {r}: status {info.status}")
            t_req += info.seconds
            iters.append(info.iterations)
            moved += info.iterations * (2.0 * (8.0 * A.nnzb * r["bs"] ** 2 + 4.0 * A.nnzb + 4.0 * (A.n_rows + 1)
                                               + 16.0 * n) + (17 if use_jacobi else 15) * 8.0 * n)
            xr.append(x.download())
        total += t_req
        per_req.append(t_req)
        xs.append(np.stack(xr))
        S.destroy()
        A.destroy()
    return total, np.array(per_req), xs, iters, moved, gen_s


SERVE_F16_MIN_BS = 4  # the server's default inner_f16_min_bs


SERVE_SLOTS = "XRPASBC"


def request_bytes_by_slot(r, outer, inner, jacobi=False, f16_min_bs=SERVE_F16_MIN_BS):
    """Compulsory HBM bytes of one served request (right-hand sides share every
    matrix read) by slot of the batched iteration, per row and right-hand side
    unless stated.  Refinement step 0 (x = 0): X writes x' (8 B), R reads b
    and writes r32 (12 B), no matrix.  Later steps: X x' = x + d (x, d read,
    x' written: 20 B), then R reads the FP64 matrix once with x', b read and
    r32 written (20 B).  Inner iteration: the inner matrix twice (A, B) and
    76 B of FP32 vectors: P p = ri + beta (p - omega v): 16; A v = A p: p, r32
    read, v written: 12; S s = ri - alpha v: 12; B t = A s: s read, t written:
    8; C the updates: s, p, d, t, r32 read, d, ri written: 28; with Jacobi
    the FP32 inverse diagonal read once per row in A, B and C.  Requests with
    bs >= f16_min_bs keep the inner matrix in FP16 with one FP32 scale per
    block row.  Matrix index bytes: block columns and row pointers."""
    side, bs, nr = r["side"], r["bs"], r["nrhs"]
    nodes = side * side
    nnzb = (3 * side - 2) ** 2
    n = nodes * bs
    idx = 4.0 * nnzb + 4.0 * (nodes + 1)
    steps = max(outer - 1, 0)
    first = 1.0 if outer > 0 else 0.0
    inner_vals = 2.0 * nnzb * bs * bs + 4.0 * nodes if (f16_min_bs and bs >= f16_min_bs) else 4.0 * nnzb * bs * bs
    jac = 4.0 * n if jacobi else 0.0
    return {"X": first * nr * n * 8.0 + steps * nr * n * 20.0,
            "R": first * nr * n * 12.0 + steps * (8.0 * nnzb * bs * bs + idx + nr * n * 20.0),
            "P": inner * nr * n * 16.0,
            "A": inner * (inner_vals + idx + nr * n * 12.0 + jac),
            "S": inner * nr * n * 12.0,
            "B": inner * (inner_vals + idx + nr * n * 8.0 + jac),
            "C": inner * (nr * n * 28.0 + jac)}


def request_bytes_moved(r, outer, inner, jacobi=False, f16_min_bs=SERVE_F16_MIN_BS):
    """Compulsory HBM bytes of one served request: the sum over the slots of
    request_bytes_by_slot."""
    return sum(request_bytes_by_slot(r, outer, inner, jacobi, f16_min_bs).values())


def serve_replay(reqs, arrivals, lam, cfg, ref_x=None):
    """Replays the requests at `lam` requests/s in real time through
    bmfp_server.  Latency = completion seen by the host - scheduled arrival;
    solves/s = requests / (last completion - first arrival).  Like the plain
    comparator's timing, solutions stay in device memory; with ref_x (an
    untimed accuracy pass) they are downloaded and compared."""
    S = bm.Server(**cfg)
    schedule = S.schedule()
    n = len(reqs)
    arr = arrivals[:n] / lam
    lat = np.zeros(n)
    ids = {}
    xs = {}
    worst_res, outer, inner, moved = 0.0, [], [], 0.0
    by_slot = dict.fromkeys(SERVE_SLOTS, 0.0)
    i = done = 0
    t0 = time.perf_counter()
    t_last = 0.0
    next_log = 5.0
    while done < n:
        now = time.perf_counter() - t0
        while i < n and arr[i] <= now:
            r = reqs[i]
            ids[S.submit(S.stencil_request(r["side"], r["side"], r["bs"], 0, r["seed"], r["nrhs"],
                                           r["rhs_seed"], 1e-10))] = i
            i += 1
        a, q = S.step()
        while True:
            p = S.poll(want_x=ref_x is not None)
            if p is None:
                break
            res, x = p
            j = ids[res.id]
            t_last = time.perf_counter() - t0
            lat[j] = t_last - arr[j]
            if res.status != bm.OK:
                raise RuntimeError(f"served request {reqs[j]} failed: status {res.status}")
            worst_res = max(worst_res, res.rel_residual)
            outer.append(res.outer_iterations)
            inner.append(res.inner_iterations)
            for k, v in request_bytes_by_slot(reqs[j], res.outer_iterations, res.inner_iterations,
                                              bool(cfg.get("use_jacobi", 0)),
                                              cfg.get("inner_f16_min_bs", SERVE_F16_MIN_BS)).items():
                by_slot[k] += v
                moved += v
            if ref_x is not None:
                xs[j] = x
            done += 1
        if now > next_log:
            log(f"[bench]   lambda={lam}: {done}/{n} done, {i - done} in the server, t={now:.1f} s")
            next_log += 5.0
        if a == 0 and q == 0 and i < n:
            time.sleep(max(0.0, arr[i] - (time.perf_counter() - t0)))
    busy = S.stats()["busy_seconds"]
    S.destroy()
    span = t_last - arr[0]
    out = {"lambda_per_s": lam, "requests": n, "span_s": round(span, 4),
           "solves_per_s": round(n / span, 2),
           "rhs_solves_per_s": round(sum(r["nrhs"] for r in reqs) / span, 2),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "max_ms": round(float(lat.max()) * 1e3, 3),
           "gpu_busy_frac": round(busy / span, 3),
           "busy_GBs": round(moved / busy / 1e9, 1),
           "compulsory_GB": round(moved / 1e9, 2),
           "compulsory_GB_by_slot": {k: round(v / 1e9, 2) for k, v in by_slot.items()},
           "max_rel_residual": worst_res,
           "mean_outer": round(float(np.mean(outer)), 2), "mean_inner": round(float(np.mean(inner)), 1),
           "schedule": {"slot_groups": schedule[0], "xcd_run": schedule[1], "chunks_per_tile": schedule[2]}}
    if ref_x is not None:
        out["max_err_vs_fp64"] = max(float((np.linalg.norm(xs[j] - ref_x[j], axis=1) /
                                            np.linalg.norm(ref_x[j], axis=1)).max()) for j in xs)
    return out


def serving_comparison(n_requests, n_low, n_plain, lambdas, max_iter, stream, cfg_overrides=None, plain=True,
                       jacobi=True):
    """Config 4: the 20,000-request stream replayed in real time at each rate
    (the first n_low requests at rates below 500/s, a steady-state prefix),
    against the plain FP64 BiCGStab solving the first n_plain requests one at
    a time (a stated sample; its solves/s is the comparison base)."""
    reqs, arrivals = serving_stream()
    log(f"[bench] serving stream: {len(reqs)} requests; replayed {n_requests} (lambda >= 500/s) and {n_low} "
        f"(below); plain sample {n_plain}")
    cfg = {"memory_budget_bytes": int(200e9)}
    cfg.update(cfg_overrides or {})
    sample = reqs[:n_plain]
    plain_res = None
    accuracy = None
    if plain:
        t0 = time.perf_counter()
        plain_s, plain_per, ref_x, plain_it, plain_bytes, plain_gen = plain_one_at_a_time(sample, max_iter, stream)
        log(f"[bench] plain FP64 one at a time: {plain_s:.3f} s ({plain_gen:.3f} s generating) "
            f"({n_plain / plain_s:.1f} solves/s), wall {time.perf_counter() - t0:.1f} s")
        plain_res = {"solves_per_s": round(n_plain / plain_s, 2), "requests": n_plain,
                     "seconds": round(plain_s, 4), "generation_seconds": round(plain_gen, 4),
                     "compulsory_GB": round(plain_bytes / 1e9, 2), "GBs": round(plain_bytes / plain_s / 1e9, 1),
                     "p50_ms": round(float(np.percentile(plain_per, 50)) * 1e3, 3),
                     "p99_ms": round(float(np.percentile(plain_per, 99)) * 1e3, 3),
                     "mean_iterations": round(float(np.mean(plain_it)), 1),
                     "timing": "per request: device matrix generation + solve device time (info.seconds) of "
                               "each right-hand side, one request after another"}
        # the sample once more through the server, all at t = 0, solutions
        # downloaded: a request's result is bitwise the same in any mix
        acc = serve_replay(sample, arrivals, float("inf"), cfg, ref_x)
        accuracy = {"max_err_vs_fp64": acc["max_err_vs_fp64"], "max_rel_residual": acc["max_rel_residual"],
                    "reference": f"plain FP64 GPU BiCGStab solutions (1e-10) of the first {n_plain} requests",
                    "compulsory_bytes_ratio_plain_over_served": round(plain_bytes / (acc["compulsory_GB"] * 1e9), 3)}
        log(f"[bench] served vs plain FP64 on the sample: max rel err {acc['max_err_vs_fp64']:.2e}, "
            f"max rel residual {acc['max_rel_residual']:.2e}; compulsory bytes plain / served "
            f"{accuracy['compulsory_bytes_ratio_plain_over_served']:.2f}")
    runs = []
    for lam in lambdas:
        nr = n_requests if lam >= 500 else n_low
        r = serve_replay(reqs[:nr], arrivals, lam, cfg)
        r["speedup_vs_plain"] = round(r["solves_per_s"] / plain_res["solves_per_s"], 3) if plain else None
        vs = f"{r['speedup_vs_plain']}x plain" if plain else "no plain comparator"
        log(f"[bench] serve lambda={lam} ({nr} requests): {r['solves_per_s']} solves/s ({vs}), "
            f"p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms, busy {r['gpu_busy_frac']}, {r['busy_GBs']} GB/s")
        runs.append(r)
    jac = None
    if plain and jacobi:
        # point-Jacobi on both sides (the reference's driver solves with
        # gpu_bicgstab(A, true)), on the plain sample, all at t = 0
        pj_s, pj_per, pj_x, pj_it, pj_bytes, _ = plain_one_at_a_time(sample, max_iter, stream, use_jacobi=True)
        acc_j = serve_replay(sample, arrivals, float("inf"), dict(cfg, use_jacobi=1), pj_x)
        rj = serve_replay(sample, arrivals, float("inf"), dict(cfg, use_jacobi=1))
        rj["max_err_vs_fp64"] = acc_j["max_err_vs_fp64"]
        burst = serve_replay(sample, arrivals, float("inf"), cfg)
        rj["speedup_vs_plain_jacobi"] = round(rj["solves_per_s"] / (n_plain / pj_s), 3)
        log(f"[bench] the sample at once: served {burst['solves_per_s']} solves/s "
            f"({burst['solves_per_s'] / plain_res['solves_per_s']:.3f}x plain); Jacobi on both sides: served "
            f"{rj['solves_per_s']} vs plain {n_plain / pj_s:.1f} solves/s ({rj['speedup_vs_plain_jacobi']}x)")
        jac = {"plain_fp64_jacobi": {"solves_per_s": round(n_plain / pj_s, 2), "seconds": round(pj_s, 4),
                                     "mean_iterations": round(float(np.mean(pj_it)), 1)},
               "served_jacobi_burst": {k: rj[k] for k in ("solves_per_s", "span_s", "p99_ms", "mean_inner",
                                                          "max_err_vs_fp64", "max_rel_residual")},
               "speedup_vs_plain_jacobi": rj["speedup_vs_plain_jacobi"],
               "served_burst_no_jacobi": {k: burst[k] for k in ("solves_per_s", "span_s", "p99_ms", "mean_inner")},
               "burst_speedup_vs_plain": round(burst["solves_per_s"] / plain_res["solves_per_s"], 3)}
    return {
        "workload": "20,000-request stream (seed 2): side U{16..512}, bs in {2,3,4,6,8}, 1-4 RHS uniform[-1,1], "
                    "tol 1e-10, diagonally dominant 9-point stencils generated on the device at admission; "
                    "served by FP64 refinement around FP32-vector BiCGStab on the FP32 (bs 2, 3) or "
                    "FP16+scale (bs 4-8) inner matrix",
        "replayed": {"lambda_ge_500": n_requests, "lambda_lt_500": n_low, "plain_sample": n_plain},
        "config": {k: v for k, v in cfg.items()},
        "plain_fp64": plain_res,
        "served": runs,
        "accuracy": accuracy,
        "jacobi": jac,
    }


def csr_shape_comparison(reps=20):
    """The CSR families (scalar, lane group, merge path) on ~1.4e8-nonzero
    matrices of different row-length shapes, FP64, x uniform[0,1]: GB/s of
    the compulsory bytes (12 nnz + 4 (n+1) + 16 n)."""
    rng = np.random.default_rng(0)
    out = []

    def run(name, rp, cols, n_cols):
        rp = np.ascontiguousarray(rp, np.int32)
        n, nnz = len(rp) - 1, int(rp[-1])
        A = bm.Matrix.csr(rp, cols, np.ones(nnz), n, n_cols)
        x = bm.DeviceArray(n_cols)
        x.fill(0, seed=1, stream_id=4)
        y = bm.DeviceArray(n)
        by = 12.0 * nnz + 4.0 * (n + 1) + 16.0 * n
        row = {"matrix": name, "n": n, "nnz": nnz, "auto": {1: "scalar", 2: "vector", 9: "merge"}[A.auto_variant()]}
        for v, vn in ((bm.CSR_SCALAR, "scalar"), (bm.CSR_VECTOR, "vector"), (bm.CSR_MERGE, "merge")):
            t = A.spmv_timed(x, y, v, warmup=3, reps=reps)
            row[vn + "_GBs"] = round(by / t / 1e9, 1)
        log(f"[bench] csr {name:26s} " + " ".join(f"{k} {row[k]}" for k in row if k.endswith("GBs"))
            + f" (auto: {row['auto']})")
        out.append(row)
        A.destroy()

    for L in (4, 16, 256):
        n = int(1.4e8 // L)
        rp = np.arange(n + 1, dtype=np.int64) * L
        cols = np.clip(np.repeat(np.arange(n, dtype=np.int64), L) + np.tile(np.arange(L) - L // 2, n) * 3,
                       0, n - 1).astype(np.int32)
        run(f"equal rows of {L}", rp, cols, n)
    for mean in (8, 40):
        n = int(1.4e8 // mean)
        L = rng.integers(1, 2 * mean, n)
        rp = np.concatenate([[0], np.cumsum(L)])
        cols = np.clip(np.repeat(np.arange(n), L) + rng.integers(-50, 50, int(rp[-1])), 0, n - 1).astype(np.int32)
        run(f"ragged rows, mean {mean}", rp, cols, n)
    n = 2_000_000
    L = np.minimum((rng.pareto(1.2, n) * 6).astype(np.int64), 500_000)
    rp = np.concatenate([[0], np.cumsum(L)])
    run("power-law rows (max 500k)", rp, rng.integers(0, n, int(rp[-1]), dtype=np.int32), n)
    return out


def parse_overrides(text):
    out = {}
    for kv in filter(None, text.split(",")):
        k, v = kv.split("=")
        out[k] = float(v) if "." in v or "e" in v else int(v)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--sizing", choices=("hbm", "literal"), default="hbm")
    ap.add_argument("--bs", default=",".join(map(str, BS_SWEEP)))
    ap.add_argument("--tune-reps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="CPU time budget for the CPU leg's sweep SpMV timings (3-10 runs per matrix)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--families", default="all", help="all | auto")
    ap.add_argument("--only", default=None,
                    help="BS:LAYOUT:FAMILY (e.g. 4:row:group_per_row): run just that kernel "
                         "--steps times (for counter collection) and print its timing")
    ap.add_argument("--copy-mb", type=float, default=2048.0,
                    help="size of the streaming-copy yardstick (0 = skip)")
    ap.add_argument("--no-solver", action="store_true",
                    help="skip the FP64 vs mixed-precision BiCGStab comparison")
    ap.add_argument("--solver-max-iter", type=int, default=20000)
    ap.add_argument("--serve-requests", type=int, default=SERVE_STREAM,
                    help="requests of the serving stream replayed at rates >= 500/s (0 = skip the serving leg)")
    ap.add_argument("--serve-low-requests", type=int, default=4000,
                    help="requests replayed at rates below 500/s (a steady-state prefix)")
    ap.add_argument("--serve-plain-sample", type=int, default=400,
                    help="requests solved one at a time by the plain FP64 BiCGStab (the comparison base)")
    ap.add_argument("--serve-no-jacobi", action="store_true", help="skip the Jacobi-on-both-sides comparison")
    ap.add_argument("--serve-lambdas", default="200,1000,5000")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--solver-only", action="store_true", help="run only the FP64 vs mixed-precision solver leg")
    ap.add_argument("--no-csr-rows", action="store_true",
                    help="skip the CSR rows (scalar / lane group / merge) of each bs in the sweep")
    ap.add_argument("--csr-only", action="store_true",
                    help="run only the CSR family comparison (scalar / lane group / merge path) over row shapes")
    ap.add_argument("--serve-no-plain", action="store_true",
                    help="skip the one-at-a-time FP64 comparator of the serving leg")
    ap.add_argument("--serve-config", default="",
                    help="server overrides, e.g. iterations_per_round=8,slot_groups=2048")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    total_values = 4.4e8 if args.sizing == "hbm" else 4.4e7
    bs_list = [int(b) for b in args.bs.split(",")]
    cpu = None
    if not (args.no_cpu or args.only or args.csr_only or args.serve_only or args.solver_only):
        cpu = start_cpu_child(bs_list, total_values, args.cpu_seconds, args.solver_max_iter)

    bm.init(0)
    dev = bm.device_info()
    log(f"[bench] device {dev['name']} CUs {dev['compute_units']} mem {dev['total_mem']/2**30:.0f} GiB")
    stream = bm.Stream()

    if args.only:
        run_only(args, total_values, stream)
        return
    if args.csr_only:
        print(json.dumps({"csr_shapes": csr_shape_comparison()}), flush=True)
        return
    if args.solver_only:
        res, _ = solver_comparison(stream, args.solver_max_iter)
        out = {"solver": res}
        if args.copy_mb > 0:                          # counter calibration
            out["copy_yardstick"] = measure_copy(args.copy_mb * 1e6, 5, stream)
        print(json.dumps(out), flush=True)
        return
    if args.serve_only:
        out = serving_comparison(args.serve_requests, args.serve_low_requests, args.serve_plain_sample,
                                 [float(v) for v in args.serve_lambdas.split(",")],
                                 args.solver_max_iter, stream, parse_overrides(args.serve_config),
                                 plain=not args.serve_no_plain, jacobi=not args.serve_no_jacobi)
        if args.copy_mb > 0:                          # counter calibration
            out["copy_yardstick"] = measure_copy(args.copy_mb * 1e6, 5, stream)
        print(json.dumps(out), flush=True)
        return

    cases = []
    sweep = {}
    for bs in bs_list:
        side = grid_side(bs, total_values)
        t0 = time.perf_counter()
        mats = build_case(bs, side, (bm.ROW_MAJOR, bm.COL_MAJOR))
        A0 = mats[bm.ROW_MAJOR]
        x = bm.DeviceArray(A0.x_len)
        x.fill(0, seed=0, stream_id=4)
        y = bm.DeviceArray(A0.y_len)
        nbytes, flops = A0.traffic()
        log(f"[bench] bs={bs} grid {side}x{side} N={A0.n_rows} nnzb={A0.nnzb} bytes={nbytes/1e9:.3f} GB "
            f"(built in {time.perf_counter()-t0:.1f} s)")
        results = {}
        best = None
        for layout, A in mats.items():
            fams = [bm.BCSR_THREAD_PER_ROW, bm.BCSR_GROUP_PER_ROW, bm.BCSR_SHUFFLE_X]
            if layout == bm.COL_MAJOR:
                fams.append(bm.BCSR_BLOCK_ROW_GROUP)
            fams.append(bm.BCSR_LIBRARY)
            if args.families == "auto":
                fams = [A.auto_variant()]
            for fam in fams:
                t = A.spmv_timed(x, y, fam, warmup=5, reps=args.tune_reps, stream=stream.handle)
                gbs = nbytes / t / 1e9
                key = f"{'row' if layout == bm.ROW_MAJOR else 'col'}/{FAMILY_NAMES[fam]}"
                results[key] = round(gbs, 1)
                log(f"[bench]   {key:28s} {t*1e6:9.1f} us  {gbs:8.1f} GB/s  {gbs/HBM_PEAK_GBS*100:5.1f}%")
                if fam != bm.BCSR_LIBRARY and (best is None or t < best[0]):
                    best = (t, layout, fam, key)
        if not args.no_csr_rows:
            # the same matrix as CSR (expanded in device memory): thread per row,
            # lane group per row and merge path, over the CSR compulsory bytes
            C = A0.to_csr()
            cbytes, _ = C.traffic()
            for fam, name in ((bm.CSR_SCALAR, "scalar"), (bm.CSR_VECTOR, "vector"), (bm.CSR_MERGE, "merge")):
                t = C.spmv_timed(x, y, fam, warmup=5, reps=args.tune_reps, stream=stream.handle)
                key = f"csr/{name}"
                results[key] = round(cbytes / t / 1e9, 1)
                log(f"[bench]   {key:28s} {t*1e6:9.1f} us  {cbytes / t / 1e9:8.1f} GB/s  "
                    f"{cbytes / t / 1e9 / HBM_PEAK_GBS * 100:5.1f}%")
            C.destroy()
        sweep[str(bs)] = results
        keep = mats[best[1]]
        for layout in list(mats):
            if layout != best[1]:
                mats[layout].destroy()
        cases.append({"bs": bs, "side": side, "A": keep, "family": best[2], "key": best[3], "x": x, "y": y,
                      "bytes": nbytes, "flops": flops})

    # ---- timed steps: every matrix once per step, best family each ----------
    def step():
        for c in cases:
            c["A"].spmv(c["x"], c["y"], c["family"], stream=stream.handle)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    e0, e1 = bm.Event(), bm.Event()
    e0.record(stream.handle)
    for _ in range(args.steps):
        step()
    e1.record(stream.handle)
    total = e1.elapsed_since(e0)
    stream.synchronize()
    step_bytes = sum(c["bytes"] for c in cases)
    ms_per_step = total / args.steps * 1e3
    value = step_bytes * args.steps / total / 1e9

    # ---- per-kernel launch time (HIP events on the work stream) -------------
    per_kernel = []
    for c in cases:
        t = c["A"].spmv_timed(c["x"], c["y"], c["family"], warmup=args.warmup, reps=args.steps,
                              stream=stream.handle)
        per_kernel.append((t, c))
    dom_t, dom = max(per_kernel, key=lambda p: p[0])
    achieved = dom["bytes"] / dom_t / 1e9
    dom_symbol = dom["A"].kernel_symbol(dom["family"])

    out = {
        "metric": "BCSR SpMV GB/s (% HBM peak); BiCGStab time to 1e-10 mixed vs FP64; p99 ms",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 9-point block stencil, x uniform[0,1])",
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "config": {
            "workload": f"FP64 BCSR SpMV block-size sweep, {args.sizing} sizing "
                        f"({total_values:.1e} stored values per matrix), best kernel family per bs",
            "bs": bs_list,
            "grid_sides": [c["side"] for c in cases],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dom_symbol, args.sizing),
            "kernel": f"bs={dom['bs']} {dom['key']}",
            "symbol": dom_symbol,
            "launch_us": round(dom_t * 1e6, 2),
            "algorithmic_bytes": dom["bytes"],
        },
        "per_bs_best": {str(c["bs"]): {"kernel": c["key"], "us": round(t * 1e6, 2),
                                         "GB/s": round(c["bytes"] / t / 1e9, 1),
                                         "frac": round(c["bytes"] / t / 1e9 / HBM_PEAK_GBS, 4)}
                        for t, c in per_kernel},
        "sweep_GBs": sweep,
    }
    if args.copy_mb > 0:
        out["copy_yardstick"] = measure_copy(args.copy_mb * 1e6, 50, stream)
    per_bs_bytes = [{"bs": c["bs"], "bytes": c["bytes"]} for c in cases]
    for c in cases:
        c["A"].destroy()
    cases.clear()
    out["system_256x256_bs2"] = config0_gpu(stream, args.steps)
    # the CPU leg runs beside the sweep only: the solver and serving legs make
    # host round trips whose latency a busy host would lengthen
    out["cpu_baseline"] = None
    x_cpu = None
    if cpu is not None:
        log("[bench] waiting for the CPU leg ...")
        t0 = time.perf_counter()
        proc, spec, host_cpus = cpu
        out["cpu_baseline"], x_cpu = cpu_baseline_result(proc, spec, host_cpus,
                                                         {c["bs"]: c["bytes"] for c in per_bs_bytes}, value,
                                                         timeout=1800)
        log(f"[bench] CPU leg done ({time.perf_counter() - t0:.1f} s waited)")
    if not args.no_solver:
        out["solver"], sols = solver_comparison(stream, args.solver_max_iter)
        if x_cpu is not None:
            # the GPU solutions against the CPU FP64 BiCGStab's
            for name, xg in sols.items():
                out["solver"][name]["err_vs_cpu_fp64"] = float(np.linalg.norm(xg - x_cpu) / np.linalg.norm(x_cpu))
    if args.serve_requests > 0:
        out["serving"] = serving_comparison(args.serve_requests, args.serve_low_requests, args.serve_plain_sample,
                                            [float(v) for v in args.serve_lambdas.split(",")],
                                            args.solver_max_iter, stream, plain=not args.serve_no_plain,
                                            jacobi=not args.serve_no_jacobi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
