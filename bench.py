#!/usr/bin/env python3
"""SPC5 beta(r,8) SpMV benchmark on one MI355X.

Workload (default): the clustered-runs matrix, N = 4,194,304 rows, lognormal
row lengths (mean 32, sigma 0.8, clipped to [1,512]), runs of 1..16 columns
within +-N/64 of the diagonal, values U[-1,1] (seed 1/2), x U[-1,1] (seed 3),
double precision, converted on the GPU to beta(1|2|4|8, 8).

A step is one y += A x over the whole matrix with every input resident in
HBM.  W untimed steps, then K timed steps between two device
synchronisations, timed with HIP events on the stream the kernel runs on.
Prints one JSON line (rank 0).

  value         GFLOP/s (2*nnz per step) of the fastest format
  roofline      algorithmic bytes per launch / mean launch time vs 8 TB/s
  cpu_baseline  the reference's OpenMP AVX-512 path (oracle/_ref) on the
                host cores, on the same matrix (16 SpMVs, as its driver does)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The host CPUs this process may use, read before any library starts an
# OpenMP runtime (one that binds its first thread would shrink the set).
HOST_CPUS = sorted(os.sched_getaffinity(0))

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV GFLOP/s + HBM GB/s (% of peak); mixed vs FP64 solve time at 1e-10"
WORKLOADS = ("clustered", "laplacian", "stencil27")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, n_override=None):
    from spc5_amd import synth
    if name == "clustered":
        N = n_override or 4_194_304
        rp, col, val = synth.clustered_runs(N, seed=1, value_seed=2)
        desc = f"clustered_runs N={N} lognormal(mean 32, sigma 0.8) runs 1-16 within +-N/64, FP64"
    elif name == "laplacian":
        n = n_override or 512
        rp, col, val = synth.laplacian2d(n)
        desc = f"laplacian2d {n}x{n} (N={n * n}), FP64"
    elif name == "stencil27":
        n = n_override or 160
        rp, col, val = synth.stencil27(n)
        desc = f"stencil27 {n}^3 (N={n ** 3}), diag 26.001, FP64"
    else:
        raise SystemExit(f"unknown workload {name}")
    return rp, col, val, desc


def time_spmv(m, x, y, steps, warmup):
    import torch
    for _ in range(warmup):
        m.spmv(x, y)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        m.spmv(x, y)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def step_times(m, x, y, steps):
    """Per-launch times (ms) of `steps` back-to-back products: an event
    before every launch and one after the last, on the launch stream.
    Returns (median, min)."""
    import statistics

    import torch
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    for i in range(steps):
        ev[i].record(stream)
        m.spmv(x, y)
    ev[steps].record(stream)
    torch.cuda.synchronize()
    t = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return statistics.median(t), min(t)


def copy_bandwidth():
    """Device-to-device copy of 2 GiB (read + write counted)."""
    import torch
    n = (2 << 30) // 8
    a = torch.empty(n, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 8 / (ms * 1e-3) / 1e9


def stream_ceiling(read_bytes, write_bytes, tiles_kib=(8, 16, 32, 64), steps=20):
    """Rate of the library's read-dominated streaming probe (each wave reads
    one contiguous tile with coalesced non-temporal loads, as the product's
    waves stream their tiles) with the byte mix of one product: the
    bandwidth a kernel that only streams these bytes reaches on this device.
    Returns (GB/s of the best tile size, {tile KiB: GB/s})."""
    import ctypes as C

    import torch
    from spc5_amd._lib import check, lib
    rb = read_bytes // 16 * 16
    wb = max(16, write_bytes // 16 * 16)
    src = torch.empty(rb // 8, dtype=torch.float64, device="cuda").uniform_(-1, 1)
    dst = torch.empty(wb // 8, dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rates = {}
    for kib in tiles_kib:
        def run():
            check(lib().spc5_hip_stream_probe(C.c_void_p(src.data_ptr()), rb, C.c_void_p(dst.data_ptr()), wb,
                                              kib * 1024, stream), "stream probe")
        t = _events_time(run, steps, 3)
        rates[kib] = round((rb + wb) / t / 1e9, 1)
    del src, dst
    torch.cuda.empty_cache()
    return max(rates.values()), rates


def access_split(m, x, steps=50):
    """Times the library's access probe on a double beta(1,8) matrix: the
    product's tiles and grid reading every header and value (the matrix
    stream), then with x gathered at every value's column in the product's
    order, then with those gathers folded into a 1024-entry window (all L1
    hits).  Splits a product's time into streaming, gathering and the rest."""
    import ctypes as C

    import torch
    from spc5_amd._lib import check, lib
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    res = {}
    for mode, name in ((0, "matrix_stream_ms"), (1, "with_x_gathers_ms"), (2, "with_l1_gathers_ms")):
        def run(mode=mode):
            check(lib().spc5_hip_access_probe(m._h, mode, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()),
                                              stream), "access probe")
        res[name] = round(_events_time(run, steps, 3) * 1e3, 5)
    stream_bytes = 8 * m.nnz + 5 * m.nblocks
    res["matrix_stream_bytes"] = stream_bytes
    res["matrix_stream_gbs"] = round(stream_bytes / (res["matrix_stream_ms"] * 1e-3) / 1e9, 1)
    return res


def traffic_from_profiles(workload, fmt):
    p = os.path.join(ROOT, "profiles", f"pmc_spmv_{workload}_{fmt}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_in_child(args, r, timeout=900):
    """Runs the CPU baseline in a fresh process (`--cpu-baseline-child`) on
    the host CPUs this process started with, its OpenMP threads bound one per
    core as the reference's job scripts do (jobs-plafrim.sh:18-19).  The
    child touches no GPU; it rebuilds the same seeded workload."""
    import subprocess
    cpus = HOST_CPUS[: args.cpu_threads] if args.cpu_threads else HOST_CPUS
    env = dict(os.environ, OMP_PROC_BIND="true", OMP_PLACES="cores", OMP_NUM_THREADS=str(len(cpus)))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", "--workload", args.workload,
           "--cpu-r", str(r), "--cpu-list", ",".join(str(c) for c in cpus)]
    if args.n:
        cmd += ["--n", str(args.n)]
    if args.cpu_max_rows:
        cmd += ["--cpu-max-rows", str(args.cpu_max_rows)]
    try:
        res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"CPU baseline exceeded {timeout} s"}
    if res.returncode != 0:
        return {"error": f"CPU baseline child exited with {res.returncode}: {res.stderr.strip()[-400:]}"}
    return json.loads(res.stdout.strip().splitlines()[-1])


def cpu_baseline(rp, col, val, r, ncols, nthreads, nloops=16, nsamples=5, max_rows=None):
    """Reference OpenMP path (compiled reference header) on the host cores:
    one split, then `nsamples` samples of `nloops` products each (the
    reference driver's loop, examples/load_mm_and_compare.cpp:656-729);
    reports the median sample's rate, with the best and worst."""
    import statistics

    import numpy as np
    try:
        import oracle as O
    except Exception as e:  # noqa: BLE001
        return {"error": f"oracle import failed: {e}"}
    if not O.ref_available():
        return {"error": "oracle/_ref/libspc5_ref.so not built"}
    n = len(rp) - 1
    if max_rows and max_rows < n:
        n = max_rows
        rp = rp[: n + 1]
        col = col[: rp[-1]]
        val = val[: rp[-1]]
    nnz = int(rp[-1])
    from spc5_amd import synth
    x = np.zeros(ncols + 15)
    x[:ncols] = synth.uniform(ncols, 3)
    y = np.zeros(n + 7)
    th, tconv, samples = O.ref_omp_bench(r, rp, col, val, x, y, nloops, nthreads, ncols=ncols, nsamples=nsamples)
    rates = [2.0 * nnz * nloops / t / 1e9 for t in samples]
    med = statistics.median(rates)
    return {"value": round(med, 3), "unit": "GFLOP/s", "cores": th, "kind": "reference",
            "sample": f"{n} rows / {nnz} values of the same matrix, beta({r},8) FP64, "
                      f"{nsamples} samples of SPC5_Spmv_omp x{nloops} after one split (conversion {tconv:.1f}s "
                      f"untimed); value = median sample",
            "min_gflops": round(min(rates), 3), "max_gflops": round(max(rates), 3),
            "samples_gflops": [round(v, 3) for v in rates],
            "ms_per_spmv": round(statistics.median(samples) / nloops * 1e3, 3),
            "cpu_model": cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def solve_benchmark(n=160, rtol=1e-10):
    """Precision form on the 27-point stencil n^3 (diag 26.001, off -1, SPD):
    wall time to relative residual rtol of double CG and of the two mixed-
    precision solvers (restarted IR-CG and reliable-update CG) with float and
    bfloat16 values, plus errors against the double
    solution and of one reduced-precision product against the double one.
    Every matrix (the double one included) uses the block shape r in
    {1, 2, 4} and row-sort setting whose product is fastest."""
    import numpy as np
    import torch
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix
    from spc5_amd import autotune
    from spc5_amd.solver import cg, ir_cg, lowp, mixed_cg

    rp, col, val = synth.stencil27(n)
    N = len(rp) - 1
    d_csr = [torch.from_numpy(a).cuda() for a in (rp, col, val)]
    shapes = {rr: DeviceMatrix.from_csr(*d_csr, rr) for rr in (1, 2, 4)}
    b = torch.from_numpy(synth.uniform(N, 7)).cuda()
    bf = b.to(torch.float32)

    def fastest(cands, x, y):
        """(r, row_sort) -> matrix; returns the key of the fastest product and
        all times.  Each candidate runs with its autotuned tile size."""
        times = {}
        for (rr, srt), m in cands.items():
            m.row_sort = srt
            autotune.tune_tiles(m, reps=5)
            times[(rr, srt)] = time_spmv(m, x, y, 20, 3)[0]
        return min(times, key=times.get), times

    def keyname(k):
        return f"r={k[0]}{' sorted' if k[1] else ''}"

    y64 = torch.zeros(N, dtype=torch.float64, device="cuda")
    k64, times64 = fastest({(rr, srt): m for rr, m in shapes.items() for srt in ((False, True) if rr > 1 else (False,))},
                           bf.to(torch.float64), y64)
    A64 = shapes[k64[0]]
    A64.row_sort = k64[1]
    r = k64[0]
    ms_64 = times64[k64]
    out = {"workload": f"stencil27 {n}^3 (N={N}, nnz={len(val)}), diag 26.001, b U[-1,1] seed 7", "rtol": rtol,
           "fp64_matrix": {"r": r, "row_sort": k64[1], "tile_values": A64.tile_target(), "spmv_ms": round(ms_64, 4),
                           "spmv_ms_by_shape": {keyname(k): round(v, 4) for k, v in times64.items()}}}
    log(f"[bench] fp64 SpMV {keyname(k64)} {ms_64:.4f} ms "
        f"(candidates {', '.join(f'{keyname(k)}: {v:.4f}' for k, v in sorted(times64.items()))})")
    # warm-up of every code path on a short solve
    cg(A64, b, rtol=1e-2, maxit=16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64, st = cg(A64, b, rtol=rtol, maxit=20000)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    out["fp64_cg"] = {"seconds": round(t64, 4), "iterations": st["iterations"], "relres": st["relres"]}
    log(f"[bench] fp64 CG: {t64 * 1e3:.1f} ms, {st['iterations']} it, relres {st['relres']:.2e}")
    nx64 = float(torch.linalg.vector_norm(x64))
    # block formats the low-precision copy may use: the same shapes, values
    # converted; the fastest product wins
    for kind in ("f32", "bf16"):
        ylo = torch.zeros(N, dtype=torch.float32, device="cuda")
        cands = {(rr, srt): lowp(m, kind) for rr, m in shapes.items()
                 for srt in ((False, True) if rr > 1 else (False,))}
        klo, times = fastest(cands, bf, ylo)
        r_lo = klo[0]
        Alo = cands[klo]
        ms_lo = times[klo]
        del cands
        # one product in reduced precision against the double product
        ylo.zero_()
        Alo.spmv(bf, ylo)
        y64.zero_()
        A64.spmv(bf.to(torch.float64), y64)
        spmv_err = float((ylo.double() - y64).abs().max() / (54.0 * float(bf.abs().max())))
        log(f"[bench] {kind} SpMV beta({r_lo},8) {keyname(klo)} {ms_lo:.4f} ms vs fp64 {keyname(k64)} "
            f"{ms_64:.4f} ms (candidates {', '.join(f'{keyname(k)}: {v:.4f}' for k, v in sorted(times.items()))})")
        out[f"{kind}_matrix"] = {"r": r_lo, "row_sort": klo[1], "value_bytes": Alo.values.element_size(),
                                 "bytes_per_spmv": Alo.algorithmic_bytes(), "spmv_ms": round(ms_lo, 4),
                                 "fp64_spmv_ms": round(ms_64, 4),
                                 "spmv_ms_by_shape": {keyname(k): round(v, 4) for k, v in times.items()},
                                 "spmv_normwise_err_vs_fp64": spmv_err}
        ir_cg(A64, Alo, b, rtol=1e-2, max_outer=1, max_inner=16)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, st = ir_cg(A64, Alo, b, rtol=rtol)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out[f"ir_cg_{kind}"] = {
            "seconds": round(t, 4), "speedup_vs_fp64": round(t64 / t, 3), "inner_iterations": st["iterations"],
            "outer": st["outer"], "relres": st["relres"],
            "x_rel_diff_vs_fp64_cg": float(torch.linalg.vector_norm(x - x64)) / nx64}
        log(f"[bench] IR-CG {kind}: {t * 1e3:.1f} ms ({t64 / t:.2f}x), {st['outer']} outer / "
            f"{st['iterations']} inner, relres {st['relres']:.2e}")
        # reliable-update CG on the same pair of matrices: the defaults (float
        # vectors; a double tail for bfloat16 only), with and without a tail,
        # and with double vectors
        tail_alt = ("_tail", {"tail_rtol": 1e-6}) if kind == "f32" else ("_no_tail", {"tail_rtol": 0.0})
        for tag, kw in (("", {}), tail_alt, ("_vec64", {"vectors": "f64"})):
            mixed_cg(A64, Alo, b, rtol=1e-2, maxit=16, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, st = mixed_cg(A64, Alo, b, rtol=rtol, **kw)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            out[f"mixed_cg_{kind}{tag}"] = {
                "seconds": round(t, 4), "speedup_vs_fp64": round(t64 / t, 3), "iterations": st["iterations"],
                "double_tail_iterations": st["tail_iterations"], "reliable_updates": st["outer"],
                "relres": st["relres"], "x_rel_diff_vs_fp64_cg": float(torch.linalg.vector_norm(x - x64)) / nx64,
                "options": kw}
            log(f"[bench] mixed CG {kind}{tag}: {t * 1e3:.1f} ms ({t64 / t:.2f}x), {st['iterations']} it "
                f"({st['tail_iterations']} in the double tail) / {st['outer']} reliable updates, "
                f"relres {st['relres']:.2e}")
        del Alo
    del shapes, d_csr
    # the precision form's result: the faster mixed-precision solver (double
    # residual b - Ax and double solution update around a float CG
    # iteration) against double CG on the same GPU
    best = min((k for k in out if k.startswith(("mixed_cg_", "ir_cg_"))), key=lambda k: out[k]["seconds"])
    out["precision_form"] = {"solver": best, "seconds": out[best]["seconds"], "fp64_cg_seconds": out["fp64_cg"]["seconds"],
                             "speedup_vs_fp64": out[best]["speedup_vs_fp64"], "relres": out[best]["relres"]}
    log(f"[bench] precision form: {best} {out[best]['seconds'] * 1e3:.1f} ms vs fp64 CG "
        f"{out['fp64_cg']['seconds'] * 1e3:.1f} ms ({out[best]['speedup_vs_fp64']:.2f}x)")
    out["fp64_bytes_per_spmv"] = A64.algorithmic_bytes()
    del A64
    torch.cuda.empty_cache()
    return out


def _events_time(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps * 1e-3


def spmm_benchmark(n=160, ks=(1, 4, 16, 64), steps=10, warmup=2, rs=(2, 4, 8)):
    """Multi-right-hand-side products on the 27-point stencil n^3, double and
    float values (beta(r,8)), row-major X and Y with k columns.  The
    autotuner picks r, the column chunk kt and the record kind (one X row
    request per stored value, or per set column of a block) by timing every
    candidate; the case's time is then measured again on its own."""
    import torch
    from spc5_amd import autotune, synth
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp

    rp, col, val = synth.stencil27(n)
    N = len(rp) - 1
    nnz = len(val)
    d_rp, d_col, d_val = (torch.from_numpy(a).cuda() for a in (rp, col, val))
    mats64 = {r: DeviceMatrix.from_csr(d_rp, d_col, d_val, r) for r in rs}
    del d_rp, d_col, d_val
    out = {"workload": f"stencil27 {n}^3 (N={N}, nnz={nnz}), beta(r,8) for r in {list(rs)}, schedule chosen by the "
                       f"autotuner", "blocks": {str(r): m.nblocks for r, m in mats64.items()}, "cases": []}
    g = torch.Generator(device="cuda").manual_seed(3)
    for tag in ("f64", "f32"):
        mats = mats64 if tag == "f64" else {r: lowp(m, "f32") for r, m in mats64.items()}
        for k in ks:
            vt = next(iter(mats.values())).vec_dtype
            X = torch.empty((N, k), dtype=vt, device="cuda").uniform_(-1, 1, generator=g)
            Y = torch.zeros((N, k), dtype=vt, device="cuda")
            (r, kt, cols, x16), cand = autotune.tune_spmm(mats, k)
            A = mats[r]
            A.spmm_columns, A.spmm_x16 = cols, x16
            t = _events_time(lambda: A.spmm(X, Y, kt=kt), steps, warmup)
            a_bytes = A.algorithmic_bytes() - (A.ncols + 2 * A.nrows) * A.vec_dtype.itemsize
            B = a_bytes + k * (A.ncols + 2 * A.nrows) * A.vec_dtype.itemsize
            kinds = {(False, False): "values", (False, True): "values x16", (True, False): "columns",
                     (True, True): "columns x16"}
            case = {"dtype": tag, "k": k, "r": r, "kt": kt,
                    "records": ("per block column" if cols else "per value") +
                    (", 16-byte X requests" if x16 else ""),
                    "ms": round(t * 1e3, 4), "gflops": round(2.0 * nnz * k / t / 1e9, 1),
                    "gbs": round(B / t / 1e9, 1), "frac_of_peak": round(B / t / 1e9 / HBM_PEAK_GBS, 4), "bytes": B,
                    "candidates_ms": {f"r={cr} kt={ckt} {kinds[(cc, cx)]}": round(v * 1e3, 4)
                                      for (cr, ckt, cc, cx), v in sorted(cand.items())}}
            out["cases"].append(case)
            log(f"[bench] SpMM {tag} k={k} beta({r},8) kt={kt} {case['records']}: {t * 1e3:.3f} ms "
                f"{case['gflops']} GFLOP/s {case['gbs']} GB/s")
            del X, Y
        if tag == "f32":
            del mats
    del mats64
    torch.cuda.empty_cache()
    return out


def batch_benchmark(count=2000, fill=0.3, steps=10, warmup=2):
    """The uneven batch: `count` banded matrices (rows log-uniform in
    [1e3, 2e5], bandwidth uniform in {4..256}, 30% fill inside the band,
    diagonal + bandwidth), generated on the GPU, r per matrix from the
    autotuner's model, all products in one persistent graph launch; double
    and float (beta(r,8) with float values)."""
    import torch
    from spc5_amd import autotune, synth
    from spc5_amd.batch import ONE_WAVE_PER_TILE, Batch
    from spc5_amd.device import DeviceMatrix
    from spc5_amd.solver import lowp

    t0 = time.perf_counter()
    params = synth.uneven_batch_params(count)
    mats, rs, nnz, rows = [], {1: 0, 2: 0, 4: 0, 8: 0}, 0, 0
    for (n, bw, seed) in params:
        rp, col, val = synth.banded_gpu(n, bw, fill, seed)
        r, _ = autotune.choose_r(rp, col, val, mode="model")
        mats.append(DeviceMatrix.from_csr(rp, col, val, r))
        rs[r] += 1
        nnz += val.numel()
        rows += n
        del rp, col, val
    setup = time.perf_counter() - t0
    log(f"[bench] batch: {count} matrices, {rows} rows, {nnz} values, r histogram {rs}, setup {setup:.1f}s")
    out = {"workload": f"{count} banded matrices, rows log-uniform [1e3,2e5], bw U{{4..256}}, fill {fill}",
           "matrices": count, "rows": rows, "nnz": nnz, "r_histogram": {str(k): v for k, v in rs.items()},
           "setup_s": round(setup, 1), "cases": []}
    def float_r2_set():
        # the same matrices in float beta(2,8), converted from the CSR again
        out_m = []
        for (n, bw, seed) in params:
            rp, col, val = synth.banded_gpu(n, bw, fill, seed)
            out_m.append(DeviceMatrix.from_csr(rp, col, val.to(torch.float32), 2, vs=8))
            del rp, col, val
        return out_m

    g = torch.Generator(device="cuda").manual_seed(5)
    for tag in ("f64", "f32"):
        # matrix sets: the model's r per matrix; for float values also r = 2
        # for every matrix (the float kernels are instruction-bound, and the
        # model was fitted on double products)
        sets = {"model": mats} if tag == "f64" else {"model": [lowp(m, "f32") for m in mats], "r=2": float_r2_set()}
        xs = [torch.empty(m.ncols, dtype=torch.float64 if tag == "f64" else torch.float32,
                          device="cuda").uniform_(-1, 1, generator=g) for m in mats]
        ys = [torch.zeros(m.nrows, dtype=x.dtype, device="cuda") for m, x in zip(mats, xs)]
        # row-sorted batches (used when every matrix shares r = 2 or 4) and
        # plain ones, for every set; the fastest is the case's time
        times, infos, wave_times, nbytes, tiles, tile_times = {}, {}, {}, {}, {}, {}
        for sname, ms in sets.items():
            for srt in (False, True):
                for m in ms:
                    m.row_sort = srt
                # tile size (timed with one wave per tile) and persistent wave
                # count are the autotuner's choices (timed candidates)
                tiles[(sname, srt)], tile_times[(sname, srt)] = autotune.tune_batch_tiles(
                    ms, lambda: Batch(ms, xs, ys, waves=ONE_WAVE_PER_TILE))
                waves, wave_times[(sname, srt)] = autotune.tune_batch_waves(lambda w: Batch(ms, xs, ys, waves=w),
                                                                            reps=3)
                b = Batch(ms, xs, ys, waves=waves)
                times[(sname, srt)] = _events_time(lambda: b.spmv(graph=True), steps, warmup)
                nbytes[(sname, srt)] = b.algorithmic_bytes()
                infos[(sname, srt)] = b.info()
                b.close()
                del b
                for m in ms:
                    m.row_sort = False
        best = min(times, key=times.get)
        sname, srt = best
        ms = sets[sname]
        t = times[best]
        info = infos[best]
        B = nbytes[best]
        tiles_hist, r_hist = {}, {}
        for m in ms:
            r_hist[str(m.r)] = r_hist.get(str(m.r), 0) + 1
            # tile sizes by power-of-two range: "512-1023" holds 512 <= size < 1024
            lo = 1 << (max(1, tiles[best] or m.tile_target()).bit_length() - 1)
            key = f"{lo}-{2 * lo - 1}"
            tiles_hist[key] = tiles_hist.get(key, 0) + 1
        case = {"dtype": tag, "ms": round(t * 1e3, 3), "gflops": round(2.0 * nnz / t / 1e9, 1),
                "gbs": round(B / t / 1e9, 1), "frac_of_peak": round(B / t / 1e9 / HBM_PEAK_GBS, 4),
                "bytes": B, "tiles": info["ntiles"], "persistent_waves": info["waves"], "launch": "hipGraph",
                "waves_candidates_ms": {str(k): round(v * 1e3, 3) for k, v in wave_times[best].items()},
                "tile_values": tiles[best],
                "tile_values_candidates_ms": {str(k): round(v * 1e3, 3) for k, v in tile_times[best].items()},
                "tile_values_histogram": dict(sorted(tiles_hist.items(), key=lambda kv: int(kv[0].split("-")[0]))),
                "matrices": sname, "r_histogram": r_hist, "row_sort": srt,
                "ms_by_candidate": {f"{k[0]}{' row-sorted' if k[1] else ''}": round(v * 1e3, 3)
                                    for k, v in times.items()}}
        out["cases"].append(case)
        log(f"[bench] batch {tag}: {t * 1e3:.2f} ms {case['gflops']} GFLOP/s {case['gbs']} GB/s "
            f"(matrices: {sname}, row sort {srt}; candidates {case['ms_by_candidate']})")
        del xs, ys, ms
        if tag == "f32":
            del sets
    del mats
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="clustered", choices=WORKLOADS)
    ap.add_argument("--n", type=int, default=None, help="override the workload size")
    ap.add_argument("--formats", default="1,2,4,8")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-max-rows", type=int, default=0)
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--convert-reps", type=int, default=3, help="conversions timed per format")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming-probe ceiling")
    ap.add_argument("--no-tune-tiles", action="store_true", help="default tile sizes (no tile autotuning)")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 variant of the SpMV matrix")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solve-n", type=int, default=160)
    ap.add_argument("--no-spmm", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-count", type=int, default=2000)
    ap.add_argument("--only-extensions", action="store_true", help="skip the headline SpMV sections")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-r", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-list", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the SpMV would shard by "
              "block-row ranges with an x all-gather; see the design notes). Run with --gpus 1.",
              file=sys.stderr)
        return 2

    if args.cpu_baseline_child:
        # the CPU baseline alone, in its own process: no GPU, no torch
        cpus = [int(c) for c in args.cpu_list.split(",") if c] or HOST_CPUS
        os.sched_setaffinity(0, cpus)
        rp, col, val, _ = make_workload(args.workload, args.n)
        out = cpu_baseline(rp, col, val, args.cpu_r, len(rp) - 1, len(cpus),
                           max_rows=args.cpu_max_rows or None)
        print(json.dumps(out), flush=True)
        return 0

    import numpy as np
    import torch
    from spc5_amd import synth
    from spc5_amd.device import DeviceMatrix

    assert torch.cuda.is_available(), "needs a HIP device"
    t_all = time.perf_counter()
    if args.only_extensions:
        out = {"only_extensions": True}
        if not args.no_solve:
            out["solve"] = solve_benchmark(args.solve_n)
        if not args.no_spmm:
            out["spmm"] = spmm_benchmark(args.solve_n)
        if not args.no_batch:
            out["batch"] = batch_benchmark(args.batch_count)
        log(f"[bench] total {time.perf_counter() - t_all:.1f}s")
        print(json.dumps(out), flush=True)
        return 0
    rp, col, val, desc = make_workload(args.workload, args.n)
    n = len(rp) - 1
    nnz = int(rp[-1])
    ncols = n
    log(f"[bench] workload {desc}: nnz={nnz} ({time.perf_counter() - t_all:.1f}s)")
    d_rp = torch.from_numpy(rp).cuda()
    d_col = torch.from_numpy(col).cuda()
    d_val = torch.from_numpy(val).cuda()
    x = torch.from_numpy(synth.uniform(ncols, 3)).cuda()
    y = torch.zeros(n, dtype=torch.float64, device="cuda")

    from spc5_amd import autotune
    formats = {}
    best = None
    for r in [int(s) for s in args.formats.split(",")]:
        # conversion on the GPU (count pass, scan, fill pass, tile plan), timed
        # on the host around a synchronised call; the first call carries
        # one-time set-up, the later ones are the steady cost
        conv = []
        for rep in range(max(1, args.convert_reps)):
            if rep:
                del m
            tc = time.perf_counter()
            m = DeviceMatrix.from_csr(d_rp, d_col, d_val, r)
            torch.cuda.synchronize()
            conv.append(time.perf_counter() - tc)
        conv_s = conv[0]
        conv_ms = 1e3 * (min(conv[1:]) if len(conv) > 1 else conv[0])
        # bytes a one-pass conversion must move: rowptr, columns and values
        # read, block-row pointers, headers and values written
        conv_bytes = 4 * (n + 1) + (4 + 8) * nnz + 4 * ((n + r - 1) // r + 1) + (4 + r) * m.nblocks + 8 * nnz
        # the tile size is the autotuner's per-matrix choice (timed candidates)
        tile, tile_times = autotune.tune_tiles(m) if not args.no_tune_tiles else (0, {})
        ms, wall = time_spmv(m, x, y, args.steps, args.warmup)
        ms_med, ms_min = step_times(m, x, y, args.steps)
        B = m.algorithmic_bytes()
        gbs = B / (ms * 1e-3) / 1e9
        gflops = 2.0 * nnz / (ms * 1e-3) / 1e9
        info = m.info()
        formats[m.format_name] = {"r": r, "nblocks": m.nblocks, "ntiles": info["ntiles"],
                                  "tile_values": m.tile_target(), "tile_chosen_by": "autotune.tune_tiles",
                                  "tile_candidates_ms": {str(k): round(v * 1e3, 5) for k, v in tile_times.items()},
                                  "bytes_per_spmv": B, "ms": round(ms, 5), "ms_median": round(ms_med, 5),
                                  "ms_min": round(ms_min, 5), "gflops": round(gflops, 2),
                                  "gbs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                                  "convert_s": round(conv_s, 3), "convert_ms": round(conv_ms, 3),
                                  "convert_bytes": conv_bytes,
                                  "convert_gbs": round(conv_bytes / (conv_ms * 1e-3) / 1e9, 1),
                                  "timed_wall_s": round(wall, 3)}
        log(f"[bench] beta({r},8): {ms:.4f} ms (median {ms_med:.4f}, min {ms_min:.4f})  {gflops:.1f} GFLOP/s  "
            f"{gbs:.0f} GB/s  "
            f"({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s)  nb={m.nblocks} tiles={info['ntiles']} "
            f"convert {conv_ms:.2f} ms (first {1e3 * conv_s:.2f}) "
            f"(tile size {m.tile_target()} values; candidates "
            f"{', '.join(f'{k}: {v * 1e3:.4f}' for k, v in tile_times.items())})")
        if r == 1:
            # where the product's time goes: the same tiles with the matrix
            # stream only, with the x gathers, with L1-resident gathers
            split = access_split(m, x)
            split["product_ms"] = round(ms, 5)
            formats[m.format_name]["access_split"] = split
            log(f"[bench] beta(1,8) access split: matrix stream {split['matrix_stream_ms']:.4f} ms "
                f"({split['matrix_stream_gbs']} GB/s), + x gathers {split['with_x_gathers_ms']:.4f} ms, "
                f"+ gathers in L1 {split['with_l1_gathers_ms']:.4f} ms, product {ms:.4f} ms")
        if best is None or ms < formats[best]["ms"]:
            best = m.format_name
        del m
        torch.cuda.empty_cache()

    # FP32 variant of the same matrix: float values, reference float format beta(r,16)
    formats_f32 = {}
    if not args.no_f32:
        d_val32 = d_val.to(torch.float32)
        x32 = x.to(torch.float32)
        y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
        for r in [int(s) for s in args.formats.split(",")]:
            m = DeviceMatrix.from_csr(d_rp, d_col, d_val32, r)
            tile, tile_times = autotune.tune_tiles(m) if not args.no_tune_tiles else (0, {})
            ms, _ = time_spmv(m, x32, y32, args.steps, args.warmup)
            ms_med, ms_min = step_times(m, x32, y32, args.steps)
            B = m.algorithmic_bytes()
            gbs = B / (ms * 1e-3) / 1e9
            formats_f32[m.format_name] = {"r": r, "vs": m.vs, "nblocks": m.nblocks, "bytes_per_spmv": B,
                                          "ms": round(ms, 5), "ms_median": round(ms_med, 5),
                                          "ms_min": round(ms_min, 5),
                                          "gflops": round(2.0 * nnz / (ms * 1e-3) / 1e9, 2),
                                          "gbs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                                          "tile_values": m.tile_target(),
                                          "tile_candidates_ms": {str(k): round(v * 1e3, 4)
                                                                 for k, v in tile_times.items()}}
            log(f"[bench] FP32 beta({r},{m.vs}): {ms:.4f} ms  {gbs:.0f} GB/s  "
                f"({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s) (tile size {m.tile_target()} values)")
            del m
        del d_val32, x32, y32
        torch.cuda.empty_cache()

    bf = formats[best]
    copy_gbs = None if args.no_copy else round(copy_bandwidth(), 1)
    ceiling = None
    if not args.no_ceiling:
        # the streaming probe with the product's byte mix: everything read
        # once, y (N values) written once
        wbytes = n * 8
        c_gbs, c_rates = stream_ceiling(bf["bytes_per_spmv"] - wbytes, wbytes)
        ceiling = {"gbs": c_gbs, "by_tile_kib": {str(k): v for k, v in c_rates.items()},
                   "read_bytes": bf["bytes_per_spmv"] - wbytes, "write_bytes": wbytes,
                   "kernel": "k_stream_probe (one contiguous tile per wave, coalesced non-temporal 8-byte "
                             "loads, 3 KiB in flight per wave)"}
        log(f"[bench] streaming ceiling (read-dominated probe, product's byte mix): {c_gbs} GB/s {c_rates}")
    out = {
        "metric": METRIC,
        "value": bf["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bf["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "nrows": n, "ncols": ncols, "nnz": nnz, "format": f"beta({bf['r']},8)",
                   "step": "y += A x, inputs resident in HBM"},
        "roofline": {"bound": "hbm", "achieved": bf["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bf["frac_of_peak"],
                     "traffic": traffic_from_profiles(args.workload, best),
                     "ceiling_measured": ceiling["gbs"] if ceiling else None,
                     "frac_of_ceiling": round(bf["gbs"] / ceiling["gbs"], 4) if ceiling else None},
        "stream_ceiling": ceiling,
        "access_split": formats[best].get("access_split"),
        "formats": formats,
        "formats_f32": formats_f32,
        "copy_gbs_measured": copy_gbs,
    }
    del d_rp, d_col, d_val, x, y
    torch.cuda.empty_cache()
    if not args.no_solve:
        out["solve"] = solve_benchmark(args.solve_n)
    if not args.no_spmm:
        out["spmm"] = spmm_benchmark(args.solve_n)
    if not args.no_batch:
        out["batch"] = batch_benchmark(args.batch_count)
    if not args.no_cpu:
        log("[bench] CPU baseline (reference OpenMP path, child process)...")
        out["cpu_baseline"] = cpu_baseline_in_child(args, bf["r"])
        cb = out["cpu_baseline"]
        log(f"[bench] CPU baseline: {cb.get('value')} GFLOP/s on {cb.get('cores')} cores "
            f"({cb.get('samples_gflops') or cb.get('error')})")
    else:
        out["cpu_baseline"] = None
    log(f"[bench] total {time.perf_counter() - t_all:.1f}s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
