"""Benchmark of the scoring hot path on one MI355X.

Default workload: plain FP32-weight scoring at biobank scale -- 100,000 people
x 1,000,000 variants x 8 scores, HWE genotypes (p ~ U(0.01, 0.5), 2%
missing), weights ~ N(0, 0.01^2), effect allele A1 for a random half.  The
25 GB .bed is generated on the device (counter-based hash) and stays resident
in HBM; one step is one pass of the scoring kernels over every variant,
accumulated into the FP64 result.

Prints one JSON line: metric svs/s = people * variants * scores / second,
plus the roofline of the dominant kernel (HIP-event time on the library's
compute stream) and the CPU baseline (restated reference CPU engine on a
bounded sample of the same workload, all host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(kernel, workload, precision, alg_bytes, live_ms):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 summary
    (profiles/*.md: FETCH_SIZE + WRITE_SIZE scaled by a copy-kernel
    calibration).  Used only when the summary covers the same workload and
    algorithmic bytes and its average kernel duration is within 15% of the
    live HIP-event measurement; otherwise None."""
    import glob

    from gnomon_amd.profile_summary import read_values

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        v = read_values(path)
        if (v.get("kernel"), v.get("workload"), v.get("precision", "")) != (kernel, workload, precision):
            continue
        if "traffic_bytes_per_launch" not in v or int(float(v.get("algorithmic_bytes_per_launch", -1))) != alg_bytes:
            continue
        rel = os.path.relpath(path, ROOT)
        ms = float(v["kernel_avg_ms"])
        if abs(ms - live_ms) > 0.15 * live_ms:
            log(f"{rel}: profiled {ms:.3f} ms vs live {live_ms:.3f} ms -- stale, traffic not reported")
            return None, rel
        return int(v["traffic_bytes_per_launch"]), rel
    return None, None


def calibrate_copy(mib):
    """A plain elementwise kernel of known traffic: reads and writes mib MiB
    (torch.bitwise_not on uint8), three launches."""
    import torch

    n = mib << 20
    src = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.bitwise_not(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": n, "calibrate_copy_bytes_written": n, "launches": 3}), flush=True)


def measure_unpack(d_bed, N, V, bpv, stride, reps, workload):
    """Standalone unpack stage (gnm_unpack_device): packed rows -> one tile
    byte per person and variant, effect-allele flip on half the rows, plus
    per-person missing counts.  Algorithmic bytes = rows read + tiles written
    + counts; timed with events on the stream the kernel is launched on."""
    import ctypes

    import torch

    from gnomon_amd import _lib

    free, _ = torch.cuda.mem_get_info()
    vu = int(min(V, (free * 0.8) // N))
    if vu <= 0:
        return None
    out = torch.empty(vu * N, dtype=torch.uint8, device="cuda")
    miss = torch.zeros(N, dtype=torch.int32, device="cuda")
    flip = (torch.arange(vu, device="cuda") % 2).to(torch.uint8)
    stream = torch.cuda.current_stream()
    L = _lib.lib()

    def launch():
        _lib.check(L.gnm_unpack_device(ctypes.c_void_p(d_bed.data_ptr()), stride, vu, N,
                                       ctypes.c_void_p(flip.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(miss.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))

    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    alg = vu * bpv + vu * N + N * 8
    gbs = alg / (ms / 1e3) / 1e9
    log(f"unpack: {vu} variants, {ms:.3f} ms, {gbs:.1f} GB/s algorithmic")
    del out
    traffic, src = profiled_traffic("k_unpack", workload, "", alg, ms)
    return {"kernel": "k_unpack", "variants": vu, "people": N, "ms": round(ms, 4), "algorithmic_bytes": alg,
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic, "traffic_source": src}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain_fp32_biobank")
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x2", "bf16x3"])
    ap.add_argument("--compare", default="fp32", help="second path timed for the report ('' = none)")
    ap.add_argument("--variants", type=int, default=0, help="override the variant count")
    ap.add_argument("--cpu-sample-variants", type=int, default=30000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="check a slice against the FP64 oracle")
    ap.add_argument("--no-unpack", action="store_true", help="skip the standalone unpack-stage measurement")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB (read + write) three times; used to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the path's sharding is "
              "described in DESIGN.md and not measured here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gnomon_amd import synth
    from gnomon_amd.pipeline import ScoringContext
    from gnomon_amd.prep import csr_from_dense

    w = synth.WORKLOADS[args.workload]
    if args.variants:
        w = w.scaled(num_variants=args.variants)
    N, V, K = w.num_people, w.num_variants, w.num_scores
    bpv = w.bytes_per_variant
    stride = (bpv + 255) // 256 * 256
    log(f"workload {w.name}: N={N} V={V} K={K} bpv={bpv} stride={stride}")

    t0 = time.time()
    W, flipped, present = synth.draw_weights(w)
    csr = csr_from_dense(W, flipped, present)
    log(f"weights/CSR ready nnz={csr.nnz} ({time.time() - t0:.1f}s)")

    d_bed = torch.empty(V * stride, dtype=torch.uint8, device="cuda")
    synth.synth_rows_device(w, d_bed.data_ptr(), stride, 0, V)
    torch.cuda.synchronize()
    log(f"device .bed generated: {V * bpv / 1e9:.2f} GB ({time.time() - t0:.1f}s)")

    def run_path(precision):
        t1 = time.time()
        ctx = ScoringContext(csr, N, precision=precision, force=True)
        log(f"[{precision}] plan built ({time.time() - t1:.1f}s)")
        for _ in range(args.warmup):
            ctx.score_rows_device(d_bed.data_ptr(), stride, 0, V)
        torch.cuda.synchronize()
        main_ms = []
        tstart = time.perf_counter()
        for _ in range(args.steps):
            ctx.score_rows_device(d_bed.data_ptr(), stride, 0, V)
            m, _tot = ctx.last_timing()  # waits on this step's events
            main_ms.append(m)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - tstart
        ms_step = elapsed * 1e3 / args.steps
        log(f"[{precision}] {ms_step:.3f} ms/step, main kernel {np.mean(main_ms):.3f} ms")
        return ctx, ms_step, float(np.mean(main_ms))

    ctx, ms_step, main_ms = run_path(args.precision)
    svs = N * V * K
    value = svs / (ms_step / 1e3)

    check = None
    if args.check:
        from oracle import oracle as orc
        S, C, done = ctx.read_accumulators()
        nrow = min(V, 2000)
        rows = d_bed.view(V, stride)[:nrow, :bpv].cpu().numpy()
        # a fresh single-pass context over the first rows vs the FP64 oracle
        from gnomon_amd.prep import ScoreCsr
        off = int(csr.row_offsets[nrow])
        head = ScoreCsr(csr.weights[:off], csr.corrections[:off], csr.columns[:off],
                        csr.row_offsets[:nrow + 1], K, csr.baseline, csr.score_variant_counts)
        with ScoringContext(head, N, precision=args.precision, force=True) as c2:
            c2.score_rows(rows, 0)
            S2, C2 = c2.finish()
        S64, C64 = orc.score_fp64(rows, head, N)
        rel = (np.abs(S2 - S64).max(0) / np.abs(S64).max(0)).max()
        check = {"variants": nrow, "counts_exact": bool(np.array_equal(C2, C64)), "max_normwise_rel": float(rel)}
        log(f"check: {check}")
    ctx.close()

    compare = None
    if args.compare and args.compare != args.precision:
        c2, ms2, main2 = run_path(args.compare)
        c2.close()
        compare = {"precision": args.compare, "ms_per_step": ms2, "value": svs / (ms2 / 1e3),
                   "main_kernel_ms": main2, "speedup_of_headline": ms2 / ms_step}

    # roofline of the dominant kernel: algorithmic bytes = packed .bed +
    # one f32 weight matrix + f64 scores and u32 counts
    alg_bytes = V * bpv + V * K * 4 + N * K * (8 + 4)
    achieved = alg_bytes / (main_ms / 1e3) / 1e9
    kname = "k_score_mfma" if args.precision != "fp32" else "k_score_fp32"
    traffic, tsrc = profiled_traffic(kname, w.name, args.precision, alg_bytes, main_ms)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "kernel": kname, "kernel_ms": round(main_ms, 4), "algorithmic_bytes": alg_bytes}

    unpack = None if args.no_unpack else measure_unpack(d_bed, N, V, bpv, stride, max(3, min(args.steps, 10)), w.name)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        from gnomon_amd.prep import ScoreCsr
        vs = min(V, args.cpu_sample_variants)
        rows = d_bed.view(V, stride)[:vs, :bpv].cpu().numpy()
        off = int(csr.row_offsets[vs])
        head = ScoreCsr(csr.weights[:off], csr.corrections[:off], csr.columns[:off],
                        csr.row_offsets[:vs + 1], K, csr.baseline, csr.score_variant_counts)
        threads = min(os.cpu_count() or 1, 16)
        t1 = time.perf_counter()
        _, _, used = orc.refcpu_run(rows, head, N, threads=threads)
        dt = time.perf_counter() - t1
        cpu = {"value": N * vs * K / dt, "unit": "sample*variant*score/s", "cores": used, "kind": "port",
               "sample": f"restated reference CPU engine on {N} people x {vs} variants x {K} scores of the "
                         f"same workload ({dt:.2f}s)"}
        log(f"cpu baseline: {cpu['value']:.3e} svs/s on {used} threads")

    out = {
        "metric": "sample*variant*score/s", "value": value, "unit": "sample*variant*score/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16x2": "bf16x2->f32->f64", "bf16x3": "bf16x3->f32->f64"}[args.precision],
        "data": "synthetic (device-generated HWE genotypes, numpy weights)",
        "config": {"workload": w.name, "people": N, "variants": V, "scores": K,
                   "bed_bytes": V * bpv, "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu, "compare": compare, "check": check,
        "unpack_stage": unpack,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
