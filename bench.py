"""Benchmark of the scoring hot path on one MI355X.

Default workload (the precision configuration): 100,000 people x 1,000,000
variants x 64 scores, HWE genotypes (A2 frequency ~ U(0.01, 0.5), 2%
missing), weights ~ Laplace(0, 1e-3) with 30% exact zeros kept present, the
effect allele A1 for a random half of the variants.  The 25 GB .bed is
generated on the device (counter-based hash) and stays resident in HBM; one
step is one pass of the scoring kernels over every variant, accumulated into
the FP64 result.

Headline path: split-BF16 MFMA with four BF16 limbs (bf16x4: weights as
integer-valued BF16 digits per power-of-two block scale, exact FP32 block
sums, FP64 across blocks).  The plain FP32 path and the split-int8 path
(i8x4) are timed on the same rows for comparison.

Prints one JSON line: metric svs/s = people * variants * scores / second,
the roofline of the dominant kernel (MFMA-bound for the split paths: executed
matrix operations over HIP-event time on the library's compute stream), the
per-score error of each path against the FP64 oracle, the standalone unpack
stage against the HBM roofline, and the CPU baseline (restated reference CPU
engine on a bounded sample of the same workload, with the host's CPU facts).

Other modes: --workload plain_fp32_biobank (8 scores), --workload adjoint_16
(differentiable path: 20 gradient steps, finite-difference and FP64 gradient
checks), --workload batching_200 (200 score files in one grouped launch),
--calibrate-copy MIB (a plain copy of known size for the HBM counters).
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
# arithmetic of each path: limb type -> exact block accumulator -> cross-block reduction
DTYPE = {"fp32": "f32*{0,1,2}->f64", "bf16x2": "bf16x2->f32->f64", "bf16x3": "bf16x3->f32->f64",
         "bf16x4": "bf16x4->f32->f64",
         "i8x3": "i8x3->i32->f64", "i8x4": "i8x4->i32->f64"}
# published dense MFMA peaks of one MI355X (no sparsity)
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "int8": 5000.0}


def cpu_facts():
    """Host CPU model, logical CPU count, the CPUs this process may run on,
    and whether simultaneous multithreading is on (Linux sysfs/procfs)."""
    facts = {"nproc": os.cpu_count()}
    try:
        facts["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                facts["model"] = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        facts["smt_active"] = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    return facts


def mfma_work(precision, N, program_rows, tiles):
    """Executed matrix operations of one scoring pass of a split path: the
    kernel multiplies every (person, program row) dosage by every column of
    its 16-column tiles (weight limbs plus missing counts, as the library
    reports them), 2 operations (multiply, add) each."""
    kind = "int8" if precision.startswith("i8") else "bf16"
    return 2 * N * program_rows * tiles * 16, kind


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
        if not (v.get("kernel", "").startswith(kernel) and (v.get("workload"), v.get("precision", "")) ==
                (workload, precision)):
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
    ceil = hbm_ceilings(d_bed, min(vu * stride, 16 << 30), stream, reps)
    return {"kernel": "k_unpack", "variants": vu, "people": N, "ms": round(ms, 4), "algorithmic_bytes": alg,
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "measured_ceilings_gbs": ceil,
            # the best rate any probe moved on this device (the measured HBM
            # peak the unpack target is stated against) and the 1:4 copy's
            "frac_of_measured_peak": round(gbs / max(ceil.values()), 4) if ceil else None,
            "frac_of_expand_ceiling": round(gbs / ceil["expand_1to4"], 4) if ceil else None,
            "traffic": traffic, "traffic_source": src}


def hbm_ceilings(d_src, nbytes, stream, reps):
    """Rates of the HBM probe kernels (gnm_hbm_probe) on nbytes of resident
    data: plain 16-byte reads, the unpack stage's non-temporal 16-byte
    stores, and a 1:4 expanding copy (every source dword to 16 output
    bytes, the unpack stage's ratio), in GB/s of bytes moved (read +
    written), with events on the launch stream."""
    import ctypes

    import torch

    from gnomon_amd import _lib

    L = _lib.lib()
    nbytes = nbytes // (1 << 20) * (1 << 20)
    free, _ = torch.cuda.mem_get_info()
    nbytes = int(min(nbytes, (free * 0.8) // 4 // (1 << 20) * (1 << 20)))
    if nbytes < (64 << 20):
        return None
    dst = torch.empty(4 * nbytes, dtype=torch.uint8, device="cuda")
    out = {}
    for kind, name, moved in ((0, "read", nbytes), (1, "write", nbytes), (2, "expand_1to4", 5 * nbytes)):
        def launch():
            _lib.check(L.gnm_hbm_probe(kind, ctypes.c_void_p(d_src.data_ptr()), nbytes,
                                       ctypes.c_void_p(dst.data_ptr()), dst.numel(),
                                       ctypes.c_void_p(stream.cuda_stream)))
        launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = round(moved / (ms / 1e3) / 1e9, 1)
    del dst
    log(f"HBM probes on {nbytes / 2**30:.1f} GiB: {out} GB/s")
    return out


def _errors(S, S64):
    """Per-score error metrics against an FP64 reference [N, K]: (i) max
    relative error over people with |S64| >= 1e-3 max|S64|, (ii) normwise
    max|S - S64| / max|S64|."""
    import numpy as np

    K = S64.shape[1]
    rel, nw = np.zeros(K), np.zeros(K)
    for k in range(K):
        ref = S64[:, k]
        m = np.abs(ref).max()
        if m == 0:
            continue
        sel = np.abs(ref) >= 1e-3 * m
        rel[k] = (np.abs(S[sel, k] - ref[sel]) / np.abs(ref[sel])).max()
        nw[k] = np.abs(S[:, k] - ref).max() / m
    return rel, nw


def _avg_rule_fraction(S, C, S64, C64, counts, tol=1e-6):
    """Fraction of (person, score) cells whose AVG = sum / (variants - missing)
    meets the reference's GPU/CPU agreement rule (score/tests/gpu_tests.rs:134-137):
    |a - b| <= tol or |a - b| / max(|a|, |b|, 1) <= tol."""
    import numpy as np

    den, den64 = counts[None, :].astype(np.float64) - C, counts[None, :].astype(np.float64) - C64
    a = np.where(den > 0, S / np.maximum(den, 1), 0.0)
    b = np.where(den64 > 0, S64 / np.maximum(den64, 1), 0.0)
    d = np.abs(a - b)
    ok = (d <= tol) | (d / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0) <= tol)
    return float(ok.mean())


def measure_host_rows(d_bed, csr, N, V, K, stride, args):
    """One pass with the .bed rows in pageable host memory, fed through
    gnm_score_rows: parallel host fill of pinned double-buffered staging,
    host-to-device copies on their own stream overlapped with scoring.  The
    PCIe-inclusive rate; best of two passes, with the default sparse ingest
    (rows in which at most half of the 64-bit words hold a non-zero call sent
    as a bitmap plus those words) and with every row sent packed."""
    from gnomon_amd.pipeline import ScoringContext

    t0 = time.time()
    host = d_bed.view(V, stride).cpu().numpy()
    log(f"host rows: {V * stride / 1e9:.1f} GB copied to host memory ({time.time() - t0:.1f}s)")
    res = {}
    for label, fraction in (("sparse_ingest", 0.5), ("packed_only", 0.0)):
        times = []
        with ScoringContext(csr, N, precision=args.precision, force=True) as hc:
            hc.set_sparse_ingest(fraction)
            hc.score_rows(host[:min(V, 65536)], 0)    # allocates the staging slots
            hc.read_accumulators()
            before = hc.ingest_stats()
            for _ in range(2):
                t1 = time.perf_counter()
                hc.score_rows(host, 0)
                hc.read_accumulators()                 # waits for the last block
                times.append(time.perf_counter() - t1)
            after = hc.ingest_stats()
        t = min(times)
        st = {k: (after[k] - before[k]) // 2 for k in after}
        log(f"host rows [{label}]: {t * 1e3:.1f} ms per pass, {V * stride / t / 1e9:.1f} GB/s of rows, "
            f"{st['bytes_to_device'] / 1e9:.2f} GB to the device, {st['sparse_rows']} rows sparse")
        res[label] = {"ms_per_pass": round(t * 1e3, 2), "value": N * V * K / t, "row_bytes_per_s": V * stride / t,
                      "bytes_to_device": st["bytes_to_device"], "dense_rows": st["dense_rows"],
                      "sparse_rows": st["sparse_rows"], "stored_words": st["stored_words"]}
    del host
    out = dict(res["sparse_ingest"])
    out["unit"] = "sample*variant*score/s"
    out["packed_only"] = res["packed_only"]
    out["path"] = "rows in pageable host memory through gnm_score_rows (pinned 2-slot staging, copy stream)"
    return out


def run_precision_report(w, csr, d_bed, stride, paths, sample=256):
    """Per-score error of each path at full size against the CPU FP64 oracle
    on the first `sample` people (all variants, all scores; the oracle runs
    over slices of people on the host threads): (i) the largest relative
    error over people with |S64| >= 1e-3 max|S64|, (ii) normwise
    max|S - S64| / max|S64|, (iii) the reference's AVG agreement rule; counts
    compared exactly."""
    import numpy as np

    from gnomon_amd.pipeline import ScoringContext
    from oracle import oracle as orc

    N, V, K = w.num_people, w.num_variants, w.num_scores
    t0 = time.time()
    rows = d_bed.view(V, stride)[:, :(sample + 3) // 4].cpu().numpy()
    S64, C64 = orc.score_fp64_parallel(rows, csr, sample)
    log(f"FP64 oracle on {sample} people x {V} variants x {K} scores ({time.time() - t0:.1f}s)")
    out = {}
    for prec in paths:
        with ScoringContext(csr, N, precision=prec, force=True) as c:
            c.score_rows_device(d_bed.data_ptr(), stride, 0, V)
            S, C = c.finish()
        rel, nw = _errors(S[:sample], S64)
        out[prec] = {"max_rel_err_per_sample": float(rel.max()), "max_normwise_err": float(nw.max()),
                     "max_rel_err_per_score": [float(f"{x:.3e}") for x in rel],
                     "counts_exact": bool(np.array_equal(C[:sample], C64)),
                     "reference_avg_rule_pass_fraction": _avg_rule_fraction(S[:sample], C[:sample], S64, C64,
                                                                             csr.score_variant_counts)}
        log(f"[{prec}] vs FP64 oracle: per-sample max rel {rel.max():.3e}, normwise {nw.max():.3e}")
    return {"oracle": "FP64 (compensated sums), first %d people, all variants and scores" % sample,
            "metric_i": "max over people with |S64| >= 1e-3 max|S64| of |S - S64| / |S64|, per score",
            "paths": out}


def run_differentiation(w, args):
    """Ridge-penalised regression through the differentiable path at 50,000
    people x 500,000 variants x 16 scores: y = G w* + N(0, 1),
    L(W) = sum_k mean_p (S[p,k] - y[p])^2 + lam ||W||^2 with S = G^T W from the
    forward of --precision (split-int8 by default) and dL/dW from the adjoint
    kernel of the same arithmetic.  Times 20 gradient
    steps; checks the gradient against central finite differences on 100
    random weights and against the FP64 CPU gradient of those weights."""
    import numpy as np
    import torch

    from gnomon_amd import autograd as ag
    from gnomon_amd import synth
    from oracle import oracle as orc

    N, V, K, lam = w.num_people, w.num_variants, w.num_scores, 1e-3
    P = args.precision
    bpv = w.bytes_per_variant
    stride = (bpv + 255) // 256 * 256
    rng = np.random.default_rng(w.seed + 7)
    d_bed = torch.empty(V, stride, dtype=torch.uint8, device="cuda")
    synth.synth_rows_device(w, d_bed.data_ptr(), stride, 0, V)
    flip_np = (rng.random(V) < 0.5).astype(np.uint8)
    d_flip = torch.from_numpy(flip_np).cuda()
    w_star = torch.from_numpy(rng.normal(0, 0.02, (V, 1)).astype(np.float32)).cuda()
    Sy, _ = ag.dense_forward(w_star, d_bed, d_flip, N, "fp32")
    y = Sy[:, 0] + torch.from_numpy(rng.normal(0, 1, N)).cuda()
    W0 = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    torch.cuda.synchronize()
    log(f"differentiation: N={N} V={V} K={K}, .bed {V * bpv / 1e9:.2f} GB on device")

    def loss_fn(S, W):
        return ((S - y[:, None]) ** 2).mean(0).sum() + lam * (W.double() ** 2).sum()

    # gradient at W0: autograd, FP64 CPU, central FD
    W = torch.from_numpy(W0).cuda().requires_grad_(True)
    loss = loss_fn(ag.polygenic_score(W, d_bed, d_flip, N, P), W)
    loss.backward()
    g_auto = W.grad.detach().cpu().numpy().astype(np.float64)
    S_ref, _ = ag.dense_forward(torch.from_numpy(W0).cuda(), d_bed, d_flip, N, "fp32")
    R64 = (2.0 / N * (S_ref - y[:, None])).cpu().numpy()
    picks = rng.choice(V * K, 100, replace=False)
    pv = np.unique(picks // K)
    rows_sub = d_bed[torch.from_numpy(pv).cuda()][:, :bpv].cpu().numpy()
    g64_rows = orc.adjoint_fp64(rows_sub, N, flip_np[pv], R64)
    g64 = np.array([g64_rows[np.searchsorted(pv, i // K), i % K] for i in picks]) + 2 * lam * W0.reshape(-1)[picks]

    def loss_at(Wnp):
        Sx, _ = ag.dense_forward(torch.from_numpy(Wnp).cuda(), d_bed, d_flip, N, "fp32")
        return float(loss_fn(Sx, torch.from_numpy(Wnp)))

    h = 2.0 ** -12
    fd = np.empty(len(picks))
    t0 = time.time()
    for i, idx in enumerate(picks):
        v, k = divmod(int(idx), K)
        Wp, Wm = W0.copy(), W0.copy()
        Wp[v, k] += h
        Wm[v, k] -= h
        fd[i] = (loss_at(Wp) - loss_at(Wm)) / ((float(Wp[v, k]) - float(W0[v, k])) + (float(W0[v, k]) - float(Wm[v, k])))
    ga = g_auto.reshape(-1)[picks]
    scale = np.abs(g64).max()
    fd_err = float(np.abs(ga - fd).max() / scale)
    g64_err = float(np.abs(ga - g64).max() / scale)
    log(f"gradient check ({time.time() - t0:.1f}s FD): max |g - FD| / max|g| = {fd_err:.3e}, "
        f"vs FP64 CPU gradient {g64_err:.3e}")

    # step size 1/lambda_max of the Hessian (2/N) G G^T + 2 lam I (power iteration)
    u = torch.from_numpy(rng.normal(0, 1, (V, 1)).astype(np.float32)).cuda()
    lam_max = 0.0
    for _ in range(12):
        s, _ = ag.dense_forward(u, d_bed, d_flip, N, "fp32")
        hu = ag.dense_adjoint((2.0 / N) * s, d_bed, d_flip, V, P).double() + 2 * lam * u.double()
        lam_max = float(hu.norm() / u.double().norm())
        u = (hu / hu.norm()).float()
    Wt = torch.zeros(V, K, dtype=torch.float32, device="cuda", requires_grad=True)
    opt = torch.optim.SGD([Wt], lr=1.0 / lam_max)
    for _ in range(2):    # warm-up steps on a throw-away copy
        Wc = Wt.detach().clone().requires_grad_(True)
        loss_fn(ag.polygenic_score(Wc, d_bed, d_flip, N, P), Wc).backward()
    torch.cuda.synchronize()
    losses = []
    t0 = time.perf_counter()
    for _ in range(20):
        opt.zero_grad()
        loss = loss_fn(ag.polygenic_score(Wt, d_bed, d_flip, N, P), Wt)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / 20
    losses = [float(x) for x in losses]
    # kernel-level split of one step
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    R = torch.randn(N, K, device="cuda", dtype=torch.float32)
    ev[0].record()
    ag.dense_forward(Wt.detach(), d_bed, d_flip, N, P)
    ev[1].record()
    ag.dense_adjoint(R, d_bed, d_flip, V, P)
    ev[2].record()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    svs = N * V * K
    log(f"20 steps: {ms_step:.2f} ms/step (forward {fwd_ms:.2f} ms, adjoint {adj_ms:.2f} ms); "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    out = {
        "metric": "gradient_steps/s", "value": 1e3 / ms_step, "unit": "steps/s", "n_gpus": 1, "steps": 20,
        "warmup": 2, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[P], "adjoint_precision": ag._adjoint_precision(P, d_bed), "data": "synthetic (device-generated HWE genotypes)",
        "config": {"workload": w.name, "people": N, "variants": V, "scores": K, "lambda": lam,
                   "step": "SGD, lr = 1/lambda_max"},
        "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
        "forward_svs_per_s": svs / (fwd_ms / 1e3), "adjoint_svs_per_s": svs / (adj_ms / 1e3),
        "gradient_check": {"weights": len(picks), "max_err_vs_central_fd": fd_err, "max_err_vs_fp64_cpu": g64_err,
                           "normalised_by": "max |FP64 gradient| over the checked weights"},
        "loss_first": losses[0], "loss_last": losses[-1],
        "loss_monotone": all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:])),
    }
    print(json.dumps(out), flush=True)


def run_batching(w, args):
    """Many score files of uneven sizes over one .bed: 200 files, each a
    log-uniform number of variants in [1000, V] drawn as a random subset,
    weights ~ N(0, 0.01^2).  One grouped context scores every file in one
    launch per pass over the rows (launch shape picked by the autotuner on a
    sample block); the baseline is 200 separate single-file contexts, each
    reading only its own variants.  Work = people * sum of file sizes."""
    import numpy as np
    import torch

    from gnomon_amd import autotune, synth
    from gnomon_amd.pipeline import ScoringContext
    from gnomon_amd.prep import csr_from_dense

    N, V, K = w.num_people, w.num_variants, w.num_scores
    bpv = w.bytes_per_variant
    stride = (bpv + 255) // 256 * 256
    t0 = time.time()
    W, flipped, present = synth.draw_weights(w)
    csr = csr_from_dense(W, flipped, present)
    sizes = present.sum(0).astype(np.int64)
    work = N * int(sizes.sum())
    log(f"batching: N={N} V={V} files={K}, sizes {sizes.min()}..{sizes.max()} (sum {sizes.sum()}), "
        f"nnz={csr.nnz} ({time.time() - t0:.1f}s)")
    d_bed = torch.empty(V * stride, dtype=torch.uint8, device="cuda")
    synth.synth_rows_device(w, d_bed.data_ptr(), stride, 0, V)
    torch.cuda.synchronize()

    t1 = time.time()
    ctx = ScoringContext(csr, N, precision=args.precision, force=True)
    log(f"grouped plan built ({time.time() - t1:.1f}s)")
    tuned = autotune.tune(ctx, d_bed.data_ptr(), stride, min(V, 131072), log=log)
    ctx.score_rows_device(d_bed.data_ptr(), stride, 0, V)
    S_g, C_g, _ = ctx.read_accumulators()
    for _ in range(args.warmup):
        ctx.score_rows_device(d_bed.data_ptr(), stride, 0, V)
    torch.cuda.synchronize()
    main_ms = []
    tstart = time.perf_counter()
    for _ in range(args.steps):
        ctx.score_rows_device(d_bed.data_ptr(), stride, 0, V)
        main_ms.append(ctx.last_timing()[0])
    torch.cuda.synchronize()
    grouped_ms = (time.perf_counter() - tstart) * 1e3 / args.steps
    ctx.close()
    log(f"grouped: {grouped_ms:.2f} ms/pass (main kernel {np.mean(main_ms):.2f} ms), shape {tuned['choice']}")

    # 200 separate launches, each file its own context and launch
    sep_ms, sep_kernel_ms, worst = 0.0, 0.0, 0.0
    t2 = time.time()
    for k in range(K):
        csr_k = csr_from_dense(W[:, k:k + 1], flipped, present[:, k:k + 1])
        with ScoringContext(csr_k, N, precision=args.precision, force=True) as c:
            torch.cuda.synchronize()
            ts = time.perf_counter()
            c.score_rows_device(d_bed.data_ptr(), stride, 0, V)
            mk, _ = c.last_timing()
            sep_ms += (time.perf_counter() - ts) * 1e3
            sep_kernel_ms += mk
            S_k, C_k = c.finish()
        if not np.array_equal(C_k[:, 0], C_g[:, k]):
            raise RuntimeError(f"file {k}: missing counts differ between grouped and separate runs")
        scale = np.abs(S_k[:, 0]).max()
        if scale > 0:
            worst = max(worst, float(np.abs(S_k[:, 0] - S_g[:, k]).max() / scale))
    log(f"separate: {sep_ms:.1f} ms for {K} launches (kernels {sep_kernel_ms:.1f} ms), "
        f"max normwise difference to grouped {worst:.2e} ({time.time() - t2:.1f}s)")

    check = None
    if not args.no_report:   # the oracle check of the grouped result runs by default
        from oracle import oracle as orc
        from gnomon_amd.prep import ScoreCsr
        sample = 128
        rows = d_bed.view(V, stride)[:, :sample // 4].cpu().numpy()
        picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), 0]
        errs = []
        for k in picks:
            csr_k = csr_from_dense(W[:, k:k + 1], flipped, present[:, k:k + 1])
            S64, C64 = orc.score_fp64_parallel(rows, csr_k, sample)
            assert np.array_equal(C64[:, 0], C_g[:sample, k])
            errs.append(float(np.abs(S_g[:sample, k] - S64[:, 0]).max() / max(np.abs(S64).max(), 1e-300)))
        check = {"files": picks, "people": sample, "max_normwise_err_vs_fp64": max(errs), "counts_exact": True}
        log(f"check vs FP64 oracle: {check}")

    out = {
        "metric": "sample*variant*score/s", "value": work / (grouped_ms / 1e3), "unit": "sample*variant*score/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": grouped_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[args.precision],
        "data": "synthetic (device-generated HWE genotypes, numpy weights and file subsets)",
        "config": {"workload": w.name, "people": N, "variants": V, "files": K,
                   "file_variants_min": int(sizes.min()), "file_variants_max": int(sizes.max()),
                   "file_variants_sum": int(sizes.sum()), "precision": args.precision},
        "grouped": {"ms_per_pass": grouped_ms, "main_kernel_ms": float(np.mean(main_ms)), "autotune": tuned},
        "separate": {"launches": K, "total_ms": sep_ms, "kernel_ms": sep_kernel_ms,
                     "speedup_of_grouped": sep_ms / grouped_ms,
                     "max_normwise_diff_to_grouped": worst},
        "check": check,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(d_bed, csr, N, V, K, bpv, stride, target_s=15.0):
    """The restated reference CPU engine on a bounded sample of the same
    workload: a short first pass estimates its rate, then a sample sized for
    about target_s seconds is timed.  Threads: the CPUs this process may run
    on, at most 16 (the host share of one GPU on the measurement boxes)."""
    from gnomon_amd.prep import ScoreCsr
    from oracle import oracle as orc

    facts = cpu_facts()
    threads = min(facts.get("affinity_cpus", facts["nproc"] or 1), 16)

    def head(vs):
        off = int(csr.row_offsets[vs])
        return ScoreCsr(csr.weights[:off], csr.corrections[:off], csr.columns[:off],
                        csr.row_offsets[:vs + 1], K, csr.baseline, csr.score_variant_counts)

    vs = min(V, 2000)
    rows = d_bed.view(V, stride)[:vs, :bpv].cpu().numpy()
    t1 = time.perf_counter()
    orc.refcpu_run(rows, head(vs), N, threads=threads)
    rate = N * vs * K / (time.perf_counter() - t1)
    # the first, short pass carries start-up costs and under-estimates the
    # rate: a sample that ran under half the target is resized and run again
    for _ in range(3):
        vs = int(min(V, max(2000, rate * target_s / (N * K))))
        rows = d_bed.view(V, stride)[:vs, :bpv].cpu().numpy()
        t1 = time.perf_counter()
        _, _, used = orc.refcpu_run(rows, head(vs), N, threads=threads)
        dt = time.perf_counter() - t1
        rate = N * vs * K / dt
        if dt >= 0.5 * target_s or vs == V:
            break
    out = {"value": N * vs * K / dt, "unit": "sample*variant*score/s", "cores": used, "kind": "port",
           "sample": f"restated reference CPU engine (oracle/gnomon_oracle.c) on {N} people x {vs} variants x "
                     f"{K} scores of the same workload, {dt:.1f} s",
           "host": facts}
    log(f"cpu baseline: {out['value']:.3e} svs/s on {used} threads ({facts})")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="precision_64")
    ap.add_argument("--precision", default="bf16x4",
                    choices=["fp32", "bf16x2", "bf16x3", "bf16x4", "i8x3", "i8x4"])
    ap.add_argument("--compare", default="fp32,i8x4",
                    help="comma-separated further paths timed on the same rows for the report ('' = none)")
    ap.add_argument("--variants", type=int, default=0, help="override the variant count")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="check a slice against the FP64 oracle")
    ap.add_argument("--no-unpack", action="store_true", help="skip the standalone unpack-stage measurement")
    ap.add_argument("--no-report", action="store_true", help="skip the per-score precision report")
    ap.add_argument("--host-rows", action="store_true",
                    help="also time a pass with the rows in host memory (PCIe-inclusive, not the headline value)")
    ap.add_argument("--launch", default="", metavar="TP,BS",
                    help="launch shape of the headline path's MFMA kernel (tiles per pass, batch stages) instead "
                         "of the default")
    ap.add_argument("--unpack-only", action="store_true",
                    help="only measure the standalone unpack stage and the HBM probe kernels on the workload's rows")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB (read + write) three times; used to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}): this build scores on one "
              "device, and splitting people over several GPUs is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gnomon_amd import synth
    from gnomon_amd.pipeline import ScoringContext
    from gnomon_amd.prep import csr_from_dense

    w = synth.WORKLOADS[args.workload]
    if args.variants:
        w = w.scaled(num_variants=args.variants)
    if w.name == "adjoint_16":
        run_differentiation(w, args)
        return
    if w.files:
        run_batching(w, args)
        return
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

    if args.unpack_only:
        unpack = measure_unpack(d_bed, N, V, bpv, stride, max(3, min(args.steps, 10)), w.name)
        print(json.dumps({"workload": w.name, "unpack_stage": unpack}), flush=True)
        return

    def run_path(precision):
        t1 = time.time()
        ctx = ScoringContext(csr, N, precision=precision, force=True)
        if args.launch and precision == args.precision:
            ctx.set_launch_config(*(int(x) for x in args.launch.split(",")))
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
    prog_rows, mfma_tiles = ctx.launch_shape()
    svs = N * V * K
    value = svs / (ms_step / 1e3)

    check = None
    if args.check:
        from oracle import oracle as orc
        from gnomon_amd.prep import ScoreCsr
        nrow = min(V, 2000)
        rows = d_bed.view(V, stride)[:nrow, :bpv].cpu().numpy()
        # a fresh single-pass context over the first rows vs the FP64 oracle
        off = int(csr.row_offsets[nrow])
        head = ScoreCsr(csr.weights[:off], csr.corrections[:off], csr.columns[:off],
                        csr.row_offsets[:nrow + 1], K, csr.baseline, csr.score_variant_counts)
        with ScoringContext(head, N, precision=args.precision, force=True) as c2:
            c2.score_rows(rows, 0)
            S2, C2 = c2.finish()
        S64, C64 = orc.score_fp64_parallel(rows, head, N)
        rel = (np.abs(S2 - S64).max(0) / np.abs(S64).max(0)).max()
        check = {"variants": nrow, "counts_exact": bool(np.array_equal(C2, C64)), "max_normwise_rel": float(rel)}
        log(f"check: {check}")
    ctx.close()

    compare = []
    for prec in [p for p in args.compare.split(",") if p and p != args.precision]:
        c2, ms2, main2 = run_path(prec)
        c2.close()
        compare.append({"precision": prec, "ms_per_step": ms2, "value": svs / (ms2 / 1e3),
                        "main_kernel_ms": main2, "headline_speedup": ms2 / ms_step})

    # algorithmic HBM bytes of one pass: packed .bed + one f32 weight matrix
    # + f64 scores and u32 counts
    alg_bytes = V * bpv + V * K * 4 + N * K * (8 + 4)
    hbm_gbs = alg_bytes / (main_ms / 1e3) / 1e9
    kname = {"fp32": "k_score_fp32"}.get(args.precision, "k_score_mx")
    traffic, tsrc = profiled_traffic(kname, w.name, args.precision, alg_bytes, main_ms)
    if args.precision == "fp32":
        roofline = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
    else:
        ops, kind = mfma_work(args.precision, N, prog_rows, mfma_tiles)
        tiles = mfma_tiles
        tflops = ops / (main_ms / 1e3) / 1e12
        peak = MFMA_PEAK_TFLOPS[kind]
        roofline = {"bound": "mfma", "achieved": round(tflops, 1), "peak": peak,
                    "unit": "TFLOP/s" if kind == "bf16" else "TOP/s", "frac": round(tflops / peak, 4),
                    "mfma": kind, "executed_ops": ops, "tiles": tiles,
                    "hbm_achieved_gbs": round(hbm_gbs, 1), "hbm_frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
    roofline.update({"traffic": traffic, "traffic_source": tsrc, "kernel": kname, "kernel_ms": round(main_ms, 4),
                     "algorithmic_bytes": alg_bytes})

    precision_report = None
    if not args.no_report:
        paths = [args.precision] + [c["precision"] for c in compare]
        if w.name == "precision_64":
            paths += [p for p in ("bf16x3",) if p not in paths]
        precision_report = run_precision_report(w, csr, d_bed, stride, paths)
    unpack = None if args.no_unpack else measure_unpack(d_bed, N, V, bpv, stride, max(3, min(args.steps, 10)),
                                                        w.name)

    host_rows = None
    if args.host_rows:
        host_rows = measure_host_rows(d_bed, csr, N, V, K, stride, args)

    cpu = None if args.no_cpu else cpu_baseline(d_bed, csr, N, V, K, bpv, stride, args.cpu_seconds)

    fp32 = next((c for c in compare if c["precision"] == "fp32"), None)
    headline = {"path": args.precision, "vs_fp32_path": round(fp32["ms_per_step"] / ms_step, 3) if fp32 else None}
    if precision_report:
        headline["max_rel_err_per_sample"] = precision_report["paths"][args.precision]["max_rel_err_per_sample"]
    out = {
        "metric": "sample*variant*score/s", "value": value, "unit": "sample*variant*score/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (device-generated HWE genotypes, numpy weights)",
        "config": {"workload": w.name, "people": N, "variants": V, "scores": K,
                   "bed_bytes": V * bpv, "precision": args.precision},
        "headline": headline, "roofline": roofline, "cpu_baseline": cpu, "compare": compare, "check": check,
        "unpack_stage": unpack, "precision_report": precision_report, "host_rows": host_rows,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
