#!/usr/bin/env python3
"""Benchmark of the skinny SGEMM/SGEMV path on one MI355X.

Workload (one "step" = one pass over the plain-FP32 shape sweep, inputs
resident in HBM):
  SGEMV y = A x and y = A^T x   for M = N = 16384 and M = 1,048,576, N = 64
  SGEMM C = A B (row-major)     for M = K = 8192, N in {1,2,4,8,16,32,64,128}
All entries uniform[-1,1] from fixed seeds, alpha = 1, beta = 0.  Every call
reads a different copy of its big operand than the call before it, so no
call re-reads data still held in the 256 MB MALL (cold-HBM numbers).

value = total algorithmic bytes of a step / step time (GB/s), with
bytes = 4(MK + KN + MN) per SGEMM and 4(MN + M + N) per SGEMV (each operand
once).  The roofline object reports the kernel that moves the most bytes
per step (SGEMV 16384 x 16384, A x) timed with HIP events around its
launches inside the timed steps; cpu_baseline times the restated reference
CPU path (single thread) on a row sample of every shape.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

SWEEP_GEMV = [(16384, 16384, False), (16384, 16384, True),
              (1 << 20, 64, False), (1 << 20, 64, True)]
SWEEP_GEMM = [(8192, n, 8192) for n in (1, 2, 4, 8, 16, 32, 64, 128)]


def dominant_kernel_label(hml, shp):
    """The kernels of one SGEMV A x call of shape shp = (M, N, trans) under
    the selected variant: the DOT streaming kernel, plus the split
    reduction when the row split spans workgroups."""
    v, _ = hml.sgemv_select(shp[0], shp[1], trans=shp[2])
    fam, lpr, rb, split, wpb = v.astuple()
    name = "k_gemv_dot_stream<%d,%d>" % (lpr, rb) if fam == 1 else "GEMV family %d" % fam
    if split > wpb:
        name += " + k_reduce_wide (row split %d over %d-wave workgroups)" % (split, wpb)
    return "%s: one SGEMV A x call, %d x %d" % (name, shp[0], shp[1])


def gemv_bytes(M, N):
    return 4.0 * (M * N + M + N)


def gemm_bytes(M, N, K):
    return 4.0 * (M * K + K * N + M * N)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the baseline sample")
    ap.add_argument("--no-tables", action="store_true",
                    help="do not load the shipped MI355X autotuned tables (selection is then "
                         "the learned selector, else the heuristic)")
    ap.add_argument("--quick", action="store_true",
                    help="only the shape sweep: leave out the reference-shape, dense "
                         "split-precision, grouped, k-means and host-pointer suites of the "
                         "JSON line")
    ap.add_argument("--dense", action="store_true",
                    help="also print the split-precision dense suite as a JSON line of its own")
    ap.add_argument("--grouped", action="store_true",
                    help="also print the grouped 512-problem suite as a JSON line of its own")
    ap.add_argument("--host", action="store_true",
                    help="with --quick or --only-extra: the host-pointer entry points "
                         "(PCIe copies included) as a JSON line of their own")
    ap.add_argument("--kmeans", action="store_true",
                    help="with --quick or --only-extra: the k-means assignment suite as a "
                         "JSON line of its own")
    ap.add_argument("--only-extra", action="store_true",
                    help="skip the main sweep, run only the requested extra suites")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the sweep, run 3 device copies and 3 fills of 256 MiB "
                         "(the byte-counter calibration kernels of a --pmc run)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU work: turn two rocprofv3 --pmc counter_collection.csv "
                         "files (FETCH_SIZE, WRITE_SIZE) of a --calibrate-copy run into "
                         "profiles/traffic_gemv_dot_16384.json")
    args = ap.parse_args()
    if args.traffic:
        traffic_from_pmc(*args.traffic)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, here)
    import hiperml_amd as hml

    hml.check(hml.lib.hmlInit())
    if args.only_extra:
        if args.dense:
            print(json.dumps(bench_dense(hml)))
        if args.grouped:
            print(json.dumps(bench_grouped(hml)))
        if args.kmeans:
            print(json.dumps(bench_kmeans(hml, cpu=not args.no_cpu)))
        if args.host:
            print(json.dumps(bench_host(hml)))
        return
    stream = hml.Stream()
    S = stream.ptr
    # Selection for the sweep: the exhaustive autotuner's shipped MI355X
    # tables (a rule of the call's own shape bucket wins), then the learned
    # selector, then the heuristic.
    tables = []
    if not args.no_tables:
        for kind in ("sgemm", "sgemv"):
            path = os.path.join(here, "config", "hml_%s_mi355x.cfg" % kind)
            if os.path.exists(path):
                hml.config_load(path, kind)
                tables.append(os.path.relpath(path, here))

    # ---- inputs: two copies of every big operand, alternated per call ----
    bufs = {}

    def big(name, n, seed):
        if name not in bufs:
            b = hml.DeviceBuffer(4 * n)
            b.fill_uniform(n, seed, stream=S)
            bufs[name] = b
        return bufs[name]

    calls = []  # (label, kind, shape, fn, bytes, flops)
    copy = 0
    for (M, N, trans) in SWEEP_GEMV:
        A = big("gemvA_%d_%d_%d" % (M, N, copy % 2), M * N, 11 + copy % 2 + M % 7)
        copy += 1
        nx, ny = (M, N) if trans else (N, M)
        x = big("x_%d" % nx, nx, 21)
        y = hml.DeviceBuffer(4 * ny)
        fn = (lambda M=M, N=N, A=A, x=x, y=y, trans=trans:
              hml.sgemv_ex(M, N, A, x, y, 1.0, 0.0, trans=trans, stream=S))
        calls.append(("sgemv_%s_%dx%d" % ("T" if trans else "N", M, N), "sgemv",
                      (M, N, trans), fn, gemv_bytes(M, N), 2.0 * M * N, (A, x, y)))
    for i, (M, N, K) in enumerate(SWEEP_GEMM):
        A = big("gemmA_%d" % (i % 2), M * K, 31 + i % 2)
        B = big("B_%d_%d" % (K, N), K * N, 41 + N)
        C = hml.DeviceBuffer(4 * M * N)
        ws = hml.lib.hmlSgemmWorkspaceSize(0, 0, 0, M, N, K, None)
        W = hml.DeviceBuffer(max(ws, 256))
        fn = (lambda M=M, N=N, K=K, A=A, B=B, C=C, W=W, ws=ws:
              hml.sgemm_ex(M, N, K, A, B, C, 1.0, 0.0, workspace=W,
                           ws_bytes=max(ws, 256), stream=S))
        calls.append(("sgemm_%dx%dx%d" % (M, N, K), "sgemm", (M, N, K), fn,
                      gemm_bytes(M, N, K), 2.0 * M * N * K, (A, B, C)))
    stream.synchronize()

    dominant = 0  # SGEMV 16384^2 A x: the most bytes per launch
    step_bytes = sum(c[4] for c in calls)
    step_flops = sum(c[5] for c in calls)

    # Timed steps: events on the launch stream only at each step's start
    # (which is also the start of the dominant call) and right after the
    # dominant call, plus one at the end: every event marker in the queue
    # costs about 2 us, so per-call events would add ~3% to the step.
    ev_dom = [[hml.Event(), hml.Event()] for _ in range(args.steps)]
    ev_end = hml.Event()
    for _ in range(args.warmup):
        for c in calls:
            c[3]()
    stream.synchronize()
    hml.synchronize()

    t0 = time.perf_counter()
    for s in range(args.steps):
        ev_dom[s][0].record(S)
        for i, c in enumerate(calls):
            c[3]()
            if i == dominant:
                ev_dom[s][1].record(S)
    ev_end.record(S)
    stream.synchronize()
    hml.synchronize()
    wall = time.perf_counter() - t0
    gpu_step_ms = ev_dom[0][0].elapsed_ms(ev_end) / args.steps
    ms_per_step = max(gpu_step_ms, 1e-9)
    dom_ms = sum(e[0].elapsed_ms(e[1]) for e in ev_dom) / args.steps

    # per-call breakdown: a separate pass over the same calls with an event
    # after every call (not part of the timed steps)
    breakdown_steps = max(1, min(args.steps, 5))
    ev = [[hml.Event() for _ in range(len(calls) + 1)] for _ in range(breakdown_steps)]
    for s in range(breakdown_steps):
        ev[s][0].record(S)
        for i, c in enumerate(calls):
            c[3]()
            ev[s][i + 1].record(S)
    stream.synchronize()
    per_call = np.zeros(len(calls))
    for s in range(breakdown_steps):
        for i in range(len(calls)):
            per_call[i] += ev[s][i].elapsed_ms(ev[s][i + 1])
    per_call /= breakdown_steps

    # accuracy of the last outputs against the FP64 check kernel
    shapes = []
    for i, c in enumerate(calls):
        label, kind, shp, _, nbytes, flops, ops = c
        if kind == "sgemv":
            M, N, trans = shp
            nx, ny = (M, N) if trans else (N, M)
            st = hml.check_f64(ny, 1, nx, ops[0], ops[1], ops[2], transA=trans,
                               lda=N, ldb=1, ldc=1)
            Keff = nx
        else:
            M, N, K = shp
            st = hml.check_f64(M, N, K, ops[0], ops[1], ops[2])
            Keff = K
        ms = per_call[i]
        gbs = nbytes / (ms * 1e-3) / 1e9
        gfl = flops / (ms * 1e-3) / 1e9
        ai = flops / nbytes
        roof_gbs = min(HBM_PEAK_GBS, FP32_PEAK_TFLOPS * 1e3 / ai)
        if kind == "sgemv":
            vv, src = hml.sgemv_select(shp[0], shp[1], trans=shp[2])
        else:
            vv, src = hml.sgemm_select(*shp)
        shapes.append({"call": label, "variant": list(vv.astuple()), "selected_by": src,
                       "ms": round(ms, 5), "GB_s": round(gbs, 1),
                       "GFLOP_s": round(gfl, 1),
                       "bound": "hbm" if roof_gbs >= HBM_PEAK_GBS else "fp32",
                       "frac_of_roofline": round(gbs / roof_gbs, 4),
                       "rel_fro_err_fp64": st["rel_fro"],
                       "gammaK_bound": hml.gamma(Keff) * st["bound_fro"]})

    d = calls[dominant]
    d_ms = dom_ms
    achieved = d[4] / (d_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None,
                "kernel": dominant_kernel_label(hml, d[2]),
                "bytes_per_launch": d[4], "avg_launch_ms": round(d_ms, 5)}
    traffic_file = os.path.join(here, "profiles", "traffic_gemv_dot_16384.json")
    if os.path.exists(traffic_file):
        try:
            roofline["traffic"] = json.load(open(traffic_file)).get("traffic_bytes_per_launch")
        except Exception:
            pass

    probe = bandwidth_probe(hml, bufs["gemmA_0"], S,
                            read_src=[bufs["gemvA_16384_16384_0"], bufs["gemvA_16384_16384_1"]])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hml, calls, args.cpu_seconds)

    if args.calibrate_copy:
        n = 256 << 20
        src = bufs["gemmA_0"]  # 8192 x 8192 floats = 256 MiB
        dst = hml.DeviceBuffer(n)
        for _ in range(3):
            hml.check(hml.lib.hmlMemcpyDtoD(dst.ptr, src.ptr, n))
        for i in range(3):
            dst.fill_uniform(n // 4, 99 + i, stream=S)
        stream.synchronize()
        dst.free()
        print("calibration: 3 copies + 3 fills of %d bytes" % n, file=sys.stderr)

    extras = {}
    if not args.quick:
        # the other workloads, measured after the timed steps
        extras["reference_shape_2048x16x2048"] = bench_reference_shape(hml, cpu=not args.no_cpu)
        for b in list(bufs.values()):
            b.free()
        bufs.clear()
        extras["dense_split_precision"] = bench_dense(hml)
        extras["grouped_512"] = bench_grouped(hml)
        extras["kmeans_assign"] = bench_kmeans(hml, cpu=not args.no_cpu)
        extras["host_pointer_entries"] = bench_host(hml)

    out = {
        "metric": "SGEMM/SGEMV GFLOP/s + HBM GB/s vs plain FP32 at matched FP64-oracle error",
        "value": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic uniform[-1,1], fixed seeds, inputs resident in HBM",
        "config": {"workload": "plain FP32 shape sweep: SGEMV Ax/ATx 16384x16384 "
                               "and 1048576x64; SGEMM 8192x8192xN N=1..128 (row-major)",
                   "calls_per_step": len(calls), "alpha": 1.0, "beta": 0.0,
                   "selection": ("autotuned tables %s, then the learned selector" % tables
                                 if tables else "learned selector / heuristic")},
        "gflops": round(step_flops / (ms_per_step * 1e-3) / 1e9, 1),
        "wall_s_timed": round(wall, 4),
        "shapes_timing": "per-call events in %d extra steps after the timed ones" % breakdown_steps,
        "roofline": roofline,
        "achievable_bandwidth": probe,
        "cpu_baseline": cpu,
        "shapes": shapes,
    }
    out.update(extras)
    print(json.dumps(out))
    sys.stdout.flush()
    if args.dense and args.quick:
        print(json.dumps(bench_dense(hml)))
    if args.grouped and args.quick:
        print(json.dumps(bench_grouped(hml)))
    if args.kmeans and args.quick:
        print(json.dumps(bench_kmeans(hml, cpu=not args.no_cpu)))
    if args.host and args.quick:
        print(json.dumps(bench_host(hml)))


def bandwidth_probe(hml, src, S, nbytes=256 << 20, reps=5, read_src=None):
    """Achievable HBM bandwidth beside the vendor peak: a read-only stream
    over read_src (hmlReadProbe: 16-byte non-temporal loads, no writes; 1 GiB
    alternating between two buffers so the 256 MB MALL holds neither), a
    device-to-device copy of 256 MiB (hipMemcpy's copy kernel: reads and
    writes nbytes) and the library's fill kernel over 256 MiB (writes only),
    HIP-event timed."""
    read = None
    if read_src:
        sink = hml.DeviceBuffer(256)
        rb = read_src[0].nbytes // 16 * 16
        for b in read_src:
            hml.check(hml.lib.hmlReadProbe(b.ptr, rb, sink.ptr, S))
        e0, e1 = hml.Event(), hml.Event()
        e0.record(S)
        for i in range(2 * reps):
            hml.check(hml.lib.hmlReadProbe(read_src[i % len(read_src)].ptr, rb, sink.ptr, S))
        e1.record(S)
        e1.synchronize()
        read = 2.0 * reps * rb / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    dst = hml.DeviceBuffer(nbytes)
    hml.check(hml.lib.hmlMemcpyDtoD(dst.ptr, src.ptr, nbytes))
    hml.synchronize()
    e0, e1 = hml.Event(), hml.Event()
    e0.record(None)
    for _ in range(reps):
        hml.check(hml.lib.hmlMemcpyDtoD(dst.ptr, src.ptr, nbytes))
    e1.record(None)
    e1.synchronize()
    copy = 2.0 * nbytes * reps / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    e0.record(S)
    for i in range(reps):
        dst.fill_uniform(nbytes // 4, 7 + i, stream=S)
    e1.record(S)
    e1.synchronize()
    write = 1.0 * nbytes * reps / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    dst.free()
    out = {"copy_read_plus_write_GB_s": round(copy, 1), "fill_write_GB_s": round(write, 1),
           "bytes": nbytes, "frac_of_peak": {"copy": round(copy / HBM_PEAK_GBS, 4),
                                            "fill": round(write / HBM_PEAK_GBS, 4)}}
    if read is not None:
        out["read_only_GB_s"] = round(read, 1)
        out["read_only_bytes"] = read_src[0].nbytes // 16 * 16
        out["frac_of_peak"]["read_only"] = round(read / HBM_PEAK_GBS, 4)
    return out


def bench_reference_shape(hml, cpu=True, reps=20):
    """The reference app's plain FP32 case: row-major C = 1.0 A B + 0.0 C,
    M = K = 2048, N = 16, A and B uniform[-1,1] from a fixed seed, C
    prefilled with finite values.  GPU: hmlSgemmEx on device buffers, HIP
    events, each call after a cache flush (cold: the 17 MB of operands fit
    the 256 MB MALL) and back to back (warm).  CPU: the restated reference
    CPU path (one thread) on the full problem.  Both against the FP64 check
    kernel, plus the reference app's checksum test (relative difference of
    the sums of C, apps/gemm/hml_sgemm_main.cu:294-305,497-511)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    M, N, K = 2048, 16, 2048
    rng = np.random.default_rng(2048)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (M, N)).astype(np.float32)
    dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
    dC = hml.DeviceBuffer.from_numpy(C0)
    S = hml.Stream()
    run = lambda: hml.sgemm_ex(M, N, K, dA, dB, dC, 1.0, 0.0, stream=S.ptr)
    run()
    S.synchronize()
    cold = []
    e0, e1 = hml.Event(), hml.Event()
    for _ in range(reps):
        hml.flush_caches(S.ptr)
        e0.record(S.ptr)
        run()
        e1.record(S.ptr)
        e1.synchronize()
        cold.append(e0.elapsed_ms(e1))
    e0.record(S.ptr)
    for _ in range(reps):
        run()
    e1.record(S.ptr)
    e1.synchronize()
    warm = e0.elapsed_ms(e1) / reps
    cold_ms = float(np.median(cold))
    nbytes = 4.0 * (M * K + K * N + M * N)
    gpu = dC.to_numpy((M, N))
    st = hml.check_f64(M, N, K, dA, dB, dC)
    v, src = hml.sgemm_select(M, N, K)
    out = {"workload": "row-major SGEMM C = A B, M = K = 2048, N = 16, uniform[-1,1] (seed 2048)",
           "variant": list(v.astuple()), "gpu_cold_ms": round(cold_ms, 5),
           "gpu_warm_ms": round(warm, 5),
           "gpu_cold_GB_s": round(nbytes / cold_ms / 1e6, 1),
           "gpu_cold_frac_of_hbm_roofline": round(nbytes / cold_ms / 1e6 / HBM_PEAK_GBS, 4),
           "gpu_GFLOP_s_cold": round(2.0 * M * N * K / cold_ms / 1e6, 1),
           "gpu_rel_fro_err_fp64": st["rel_fro"]}
    if cpu:
        import oracle
        Cc = C0.copy()
        t0 = time.perf_counter()
        oracle.sgemm_cpu(Cc, A, B, 1.0, 0.0)
        dt = time.perf_counter() - t0
        R = A.astype(np.float64) @ B.astype(np.float64)
        gs, cs = float(np.sum(gpu, dtype=np.float64)), float(np.sum(Cc, dtype=np.float64))
        out.update({"cpu_path_ms_1thread": round(dt * 1e3, 2),
                    "cpu_path_GB_s": round(nbytes / dt / 1e9, 3),
                    "cpu_path_rel_fro_err_fp64": float(np.linalg.norm(Cc - R) / np.linalg.norm(R)),
                    "csum_rel_diff_gpu_vs_cpu_path": abs(gs - cs) / abs(cs),
                    "gpu_speedup_vs_cpu_path": round(dt * 1e3 / cold_ms, 1)})
    return out


def bench_dense(hml, reps=5):
    """Split-precision bf16x3 / bf16x6 / scaled f16x3 vs plain FP32 on the
    matrix cores: M = N = K = 8192 and M = N = 8192, K = 256, inputs
    uniform[-1,1] (set A) and log-uniform magnitudes in [1e-3, 1e3] with
    random sign (set B); error against the FP64 check kernel.  Each call
    includes its split pass (the split planes are part of the work)."""
    import numpy as np
    stream = hml.Stream()  # keep the object: its destructor destroys the stream
    S = stream.ptr
    rows = []
    precisions = ("fp32", "bf16x3", "bf16x6", "f16x3")
    for (M, N, K) in ((8192, 8192, 8192), (8192, 8192, 256)):
        rng = np.random.default_rng(M + K)
        for dist in ("uniform", "loguniform"):
            if dist == "uniform":
                A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
                B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
            else:
                A = (10.0 ** rng.uniform(-3, 3, (M, K)) * rng.choice([-1.0, 1.0], (M, K))).astype(np.float32)
                B = (10.0 ** rng.uniform(-3, 3, (K, N)) * rng.choice([-1.0, 1.0], (K, N))).astype(np.float32)
            dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
            dC = hml.DeviceBuffer(4 * M * N)
            del A, B
            for prec in precisions:
                ws = hml.lib.hmlSgemmDenseWorkspaceSize(hml.PRECISION[prec], M, N, K)
                W = hml.DeviceBuffer(ws)
                run = lambda: hml.sgemm_dense(M, N, K, dA, dB, dC, precision=prec,
                                              workspace=W, ws_bytes=ws, stream=S)
                run()
                hml.synchronize()
                e0, e1 = hml.Event(), hml.Event()
                e0.record(S)
                for _ in range(reps):
                    run()
                e1.record(S)
                e1.synchronize()
                ms = e0.elapsed_ms(e1) / reps
                st = hml.check_f64(M, N, K, dA, dB, dC)
                rows.append({"M": M, "N": N, "K": K, "inputs": dist, "precision": prec,
                             "ms": round(ms, 4), "TFLOP_s": round(2.0 * M * N * K / ms / 1e9, 2),
                             "rel_fro_err_fp64": st["rel_fro"],
                             "max_err_over_bound": st["max_rel_bound"]})
                del W
    for r in rows:
        base = [q for q in rows if q["M"] == r["M"] and q["K"] == r["K"] and
                q["inputs"] == r["inputs"] and q["precision"] == "fp32"][0]
        r["speedup_vs_fp32"] = round(base["ms"] / r["ms"], 3)
        r["err_vs_fp32"] = round(r["rel_fro_err_fp64"] / base["rel_fro_err_fp64"], 3)
    # matched error: the fastest split whose error is no worse than FP32's on
    # every shape and input set, and its smallest speedup over those cases
    summary = {}
    for prec in precisions[1:]:
        mine = [r for r in rows if r["precision"] == prec]
        summary[prec] = {"max_err_vs_fp32": max(r["err_vs_fp32"] for r in mine),
                         "min_speedup_vs_fp32": min(r["speedup_vs_fp32"] for r in mine),
                         "matched_error": all(r["err_vs_fp32"] <= 1.0 for r in mine)}
    return {"suite": "dense split precision vs plain FP32 MFMA (8192^3 and 8192x8192x256, "
                     "uniform and log-uniform inputs, error vs the FP64 check kernel)",
            "peaks": {"fp32_TFLOP_s": FP32_PEAK_TFLOPS, "bf16_fp16_dense_TFLOP_s": 2516.6},
            "summary": summary, "rows": rows}


def grouped_suite_shapes(count=512, seed=0):
    """The batched workload: `count` row-major problems (M, N, K) with M and
    K log-uniform in [64, 65536] and [64, 4096] (rounded), N uniform in
    1..64, drawn in that order from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        M = int(round(np.exp(rng.uniform(np.log(64), np.log(65536)))))
        K = int(round(np.exp(rng.uniform(np.log(64), np.log(4096)))))
        N = int(rng.integers(1, 65))
        out.append((M, N, K))
    return out


def bench_grouped(hml, count=512, reps=5, seed=0):
    """512 uneven skinny problems (M log-uniform [64,65536], K log-uniform
    [64,4096], N uniform in 1..64, uniform[-1,1]) in one persistent grouped
    launch (direct and hipGraph) vs 512 sequential hmlSgemmEx calls; the
    plan's kernel kinds are chosen by hmlSgemmGroupedAutotune before the
    timed runs (the default kinds' graph time is reported beside them)."""
    import numpy as np
    S = hml.Stream()
    probs, nbytes, flops = [], 0.0, 0.0
    for i, (M, N, K) in enumerate(grouped_suite_shapes(count, seed)):
        A = hml.DeviceBuffer(4 * M * K); A.fill_uniform(M * K, 3 * i + 1, stream=S.ptr)
        B = hml.DeviceBuffer(4 * K * N); B.fill_uniform(K * N, 3 * i + 2, stream=S.ptr)
        C = hml.DeviceBuffer(4 * M * N)
        probs.append((M, N, K, A, B, C))
        nbytes += 4.0 * (M * K + K * N + M * N)
        flops += 2.0 * M * N * K
    ws = max(hml.lib.hmlSgemmWorkspaceSize(0, 0, 0, p[0], p[1], p[2], None) for p in probs)
    W = hml.DeviceBuffer(ws)
    plan = hml.GroupedPlan(probs)

    def seq():
        for M, N, K, A, B, C in probs:
            hml.sgemm_ex(M, N, K, A, B, C, workspace=W, ws_bytes=ws, stream=S.ptr)

    def timed(fn):
        fn()
        S.synchronize()
        e0, e1 = hml.Event(), hml.Event()
        e0.record(S.ptr)
        for _ in range(reps):
            fn()
        e1.record(S.ptr)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    t_seq = timed(seq)
    default_kinds = plan.kinds()
    t_default = timed(lambda: plan.run(graph=True, stream=S.ptr))
    # tile shapes chosen by the plan-level autotuner (every kernel kind of
    # every class timed on these problems; C is restored afterwards)
    tune = plan.autotune(reps=3, stream=S.ptr)
    t_dir = timed(lambda: plan.run(graph=False, stream=S.ptr))
    t_graph = timed(lambda: plan.run(graph=True, stream=S.ptr))
    info = plan.info()
    # accuracy of the grouped result: worst problem against FP64
    worst = 0.0
    for M, N, K, A, B, C in probs[:64]:
        st = hml.check_f64(M, N, K, A, B, C)
        worst = max(worst, st["max_rel_bound"] / hml.gamma(K + 2))
    plan.destroy()
    # roofline of the suite: per problem the longer of its HBM time (each
    # operand once) and its FP32 time, summed
    floor_ms = sum(max(4.0 * (M * K + K * N + M * N) / (HBM_PEAK_GBS * 1e9),
                       2.0 * M * N * K / (FP32_PEAK_TFLOPS * 1e12))
                   for M, N, K, *_ in probs) * 1e3
    out = {"suite": "grouped skinny SGEMM, %d problems (M log-uniform [64,65536], K log-uniform "
                    "[64,4096], N uniform 1..64, seed %d)" % (count, seed),
           "algorithmic_GB": round(nbytes / 1e9, 3), "GFLOP": round(flops / 1e9, 1),
           "sequential_ms": round(t_seq, 4), "grouped_ms": round(t_dir, 4),
           "grouped_graph_ms": round(t_graph, 4),
           "speedup_graph_vs_sequential": round(t_seq / t_graph, 3),
           "grouped_graph_GB_s": round(nbytes / t_graph / 1e6, 1),
           "roofline_floor_ms": round(floor_ms, 4),
           "frac_of_roofline": round(floor_ms / t_graph, 4),
           "kinds": tune["kinds"],
           "autotune_class_ms": {str(c): {k: round(v, 4) for k, v in t.items()}
                                 for c, t in tune["ms"].items()},
           "default_kinds": default_kinds, "default_kinds_graph_ms": round(t_default, 4),
           "tiles": info["tiles"], "blocks": info["blocks"],
           "max_err_over_gammaK_first64": round(worst, 4)}
    for p in probs:
        for b in p[3:]:
            b.free()
    return out


def bench_kmeans(hml, reps=5, rows=1 << 24, cpu=True):
    """k-means assignment (hmlKmeansAssign) of 2^24 rows, uniform[-1,1]
    rows and centroids, for dims 16/32/64 x clusters 16/64/256.  Algorithmic
    bytes per row: 4 dims (row) + 4 (index); flops: 2 dims clusters (the
    x.c screen on the matrix cores).  The CPU column times the restated
    reference assign loop (oracle, 1 thread) on the first 20000 rows and
    counts that sample's mismatches."""
    import numpy as np
    S = hml.Stream()
    out = []
    for D in (16, 32, 64):
        R = hml.DeviceBuffer(4 * rows * D)
        R.fill_uniform(rows * D, 11 + D, stream=S.ptr)
        A = hml.DeviceBuffer(4 * rows)
        for K in (16, 64, 256):
            C = hml.DeviceBuffer(4 * K * D)
            C.fill_uniform(K * D, 5 + K, stream=S.ptr)
            hml.kmeans_assign(R, C, A, rows, D, K, stream=S.ptr)
            S.synchronize()
            e0, e1 = hml.Event(), hml.Event()
            e0.record(S.ptr)
            for _ in range(reps):
                hml.kmeans_assign(R, C, A, rows, D, K, stream=S.ptr)
            e1.record(S.ptr)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            nbytes = rows * (4.0 * D + 4.0)
            flops = 2.0 * rows * D * K
            gbs = nbytes / ms / 1e6
            tfs = flops / ms / 1e9
            t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
            t_mfma = flops / (FP32_PEAK_TFLOPS * 1e12)
            rec = {"dims": D, "clusters": K, "ms": round(ms, 4),
                   "rows_per_s": round(rows / ms * 1e3, 0), "GB_s": round(gbs, 1),
                   "TFLOP_s": round(tfs, 2),
                   "bound": "hbm" if t_hbm >= t_mfma else "mfma",
                   "frac_of_roofline": round(max(t_hbm, t_mfma) / (ms * 1e-3), 4)}
            if cpu:
                sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
                import oracle
                n = 20000
                xs = R.to_numpy((n, D))
                cs = C.to_numpy((K, D))
                t0 = time.perf_counter()
                want = oracle.kmeans_assign(xs, cs)
                dt = time.perf_counter() - t0
                got = A.to_numpy((n,), np.uint32)
                rec["cpu_rows_per_s_1thread"] = round(n / dt, 0)
                rec["cpu_sample_mismatches"] = int((got != want).sum())
            out.append(rec)
            C.free()
        R.free()
        A.free()
    return {"suite": "k-means assignment, %d rows" % rows, "results": out}


def bench_host(hml, reps=3):
    """The reference's own contract: hmlSgemm / hmlSgemv on host arrays
    (copy in, compute, copy out, synchronous), wall-clock per call, for
    SGEMM 8192 x 16 x 8192 and SGEMV 16384 x 16384; the rate counts the
    same algorithmic bytes as the device-resident sweep, so the gap to it
    is the PCIe copies."""
    rng = np.random.default_rng(0)
    out = []
    for kind, shape in (("sgemm", (8192, 16, 8192)), ("sgemv", (16384, 16384))):
        if kind == "sgemm":
            M, N, K = shape
            A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
            B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
            C = np.zeros((M, N), np.float32)
            fn = lambda: hml.sgemm(C, A, B, 1.0, 0.0)
            nbytes = 4.0 * (M * K + K * N + M * N)
        else:
            M, N = shape
            A = rng.uniform(-1, 1, (M, N)).astype(np.float32)
            x = rng.uniform(-1, 1, N).astype(np.float32)
            y = np.zeros(M, np.float32)
            fn = lambda: hml.sgemv(y, A, x, 1.0, 0.0)
            nbytes = 4.0 * (M * N + M + N)
        res = {"call": "%s %s (host pointers)" % (kind, "x".join(map(str, shape)))}
        # staged: pinned double-buffered chunks (the default); plain: one
        # hipMemcpy per operand (HML_HOST_STAGING=off)
        for mode, env in (("staged", "on"), ("plain", "off")):
            os.environ["HML_HOST_STAGING"] = env
            fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            dt = (time.perf_counter() - t0) / reps
            res[mode] = {"ms": round(dt * 1e3, 3), "GB_s": round(nbytes / dt / 1e9, 1)}
        os.environ.pop("HML_HOST_STAGING", None)
        out.append(res)
    return {"suite": "host-pointer entry points, PCIe copies included", "results": out}


CALIB_BYTES = 256 << 20


def _pmc_rows(path, counter):
    """(kernel name, grid size, counter value in bytes) per dispatch of a
    rocprofv3 counter_collection.csv (FETCH_SIZE / WRITE_SIZE are KiB)."""
    import csv
    out = []
    with open(path) as f:
        rd = csv.DictReader(f)
        cols = {c.lower(): c for c in rd.fieldnames}
        kn, cn, cv = cols["kernel_name"], cols["counter_name"], cols["counter_value"]
        gs = cols.get("grid_size") or cols.get("grid_size_x")
        for r in rd:
            if r[cn].strip() != counter:
                continue
            out.append((r[kn], int(float(r[gs])) if gs else 0, float(r[cv]) * 1024.0))
    return out


def traffic_from_pmc(fetch_csv, write_csv, dst=None):
    """HBM bytes per launch of the SGEMV 16384^2 A x kernel, each counter
    scaled by what it reads for a plain copy of CALIB_BYTES (copy:
    CALIB_BYTES read and written) and checked against the fill kernel
    (CALIB_BYTES written, nothing read)."""
    res = {}
    target = None  # (kernel, grid) of the 16384^2 launch: the largest GEMV read
    for name, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"),
                                ("write", write_csv, "WRITE_SIZE")):
        rows = _pmc_rows(path, counter)
        groups = {}
        for kn, g, v in rows:
            groups.setdefault((kn, g), []).append(v)
        gemv = [(np.mean(v), k) for k, v in groups.items() if "gemv_dot" in k[0]]
        if target is None and gemv:
            target = max(gemv)[1]
        copy = [(np.mean(v), k) for k, v in groups.items() if "copy" in k[0].lower()]
        fill = [(np.mean(v), k) for k, v in groups.items() if "fill_uniform" in k[0]]
        g_bytes = float(np.mean(groups[target])) if target in groups else None
        res[name] = {"gemv_bytes": g_bytes,
                     "gemv_kernel": target[0] if target else None,
                     "gemv_grid": target[1] if target else None,
                     "copy_bytes": max(copy)[0] if copy else None,
                     "copy_kernel": max(copy)[1][0] if copy else None,
                     "fill_bytes": max(fill)[0] if fill else None}
    f, w = res["fetch"], res["write"]
    out = {"kernel": f["gemv_kernel"], "shape": "SGEMV A x, M = N = 16384",
           "algorithmic_bytes_per_launch": gemv_bytes(16384, 16384),
           "calibration_bytes": CALIB_BYTES, "counters": res}
    if f["gemv_bytes"] and f["copy_bytes"] and w["gemv_bytes"] is not None and w["copy_bytes"]:
        kf = CALIB_BYTES / f["copy_bytes"]
        kw = CALIB_BYTES / w["copy_bytes"]
        out["fetch_factor"] = kf
        out["write_factor"] = kw
        out["traffic_bytes_per_launch"] = f["gemv_bytes"] * kf + w["gemv_bytes"] * kw
        out["traffic_over_algorithmic"] = out["traffic_bytes_per_launch"] / gemv_bytes(16384, 16384)
    if dst is None:
        here = os.path.dirname(os.path.abspath(__file__))
        dst = os.path.join(here, "profiles", "traffic_gemv_dot_16384.json")
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))
    return out


def cpu_baseline(hml, calls, budget_s):
    """Restated reference CPU path (oracle, 1 thread) on a row sample of
    every call; the full step time is extrapolated per shape from the
    sample's rows/second and value = step bytes / extrapolated time."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import oracle
    per_shape_budget = budget_s / len(calls)
    total_s = 0.0
    total_bytes = 0.0
    desc = []
    for label, kind, shp, _, nbytes, flops, ops in calls:
        if kind == "sgemv":
            M, N, trans = shp
            # reference CPU path exists for A x only; A^T x is timed as the
            # same loop over the transposed rows (same bytes and flops)
            rows_total, rowlen = (N, M) if trans else (M, N)
            est_rate = 1.0e9  # flops/s guess to size the first sample
            r = int(max(1, min(rows_total, per_shape_budget * est_rate / (2.0 * rowlen))))
            r = min(r, max(1, int(2.0e8 // (rowlen * 4))))  # <= 200 MB sample
            if trans:
                Ahost = ops[0].to_numpy((M, N))[:, :r].T.copy()
            else:
                Ahost = ops[0].to_numpy((r, N))
            x = ops[1].to_numpy((rowlen,))
            y = np.zeros(r, np.float32)
            t0 = time.perf_counter()
            oracle.sgemv_cpu(y, Ahost, x, 1.0, 0.0)
            dt = time.perf_counter() - t0
            full = dt * rows_total / r
            desc.append("%s: %d of %d rows" % (label, r, rows_total))
        else:
            M, N, K = shp
            r = int(max(1, min(M, per_shape_budget * 1.0e9 / (2.0 * N * K))))
            Ahost = ops[0].to_numpy((r, K))
            Bhost = ops[1].to_numpy((K, N))
            C = np.zeros((r, N), np.float32)
            t0 = time.perf_counter()
            oracle.sgemm_cpu(C, Ahost, Bhost, 1.0, 0.0)
            dt = time.perf_counter() - t0
            full = dt * M / r
            desc.append("%s: %d of %d rows" % (label, r, M))
        total_s += full
        total_bytes += nbytes
    return {"value": round(total_bytes / total_s / 1e9, 3), "unit": "GB/s",
            "cores": 1, "kind": "port",
            "sample": "restated hmlSgemmCpu/hmlSgemvCpu (oracle/hml_oracle.c, "
                      "-O2, 1 thread) on row samples, extrapolated per shape: " +
                      "; ".join(desc),
            "step_seconds_extrapolated": round(total_s, 2),
            "host_cpus": os.cpu_count()}


if __name__ == "__main__":
    main()
