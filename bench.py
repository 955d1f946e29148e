"""Benchmark of the batched small-GEMM path on one MI355X.

Workload (the metric's configuration): 50,000 independent SGEMMs, C = A*B
(NoTrans/NoTrans, row-major), with M, N, K each uniform on {1..128}
(splitmix64 seed 0), entries uniform in [-1, 1), all resident in HBM.  One
step = one replay of the batched plan (a hipGraph of the size-bucketed tile
kernels) over the whole batch.  FP32 MFMA, BF16X6, BF16X3 and BF16 are timed
the same way, each with its max / mean componentwise relative error
|C - C64| / (|A||B|) against an FP64 product of a sample of the GEMMs (2,000
at random plus eight for each K from 1 to 16) and its count of sampled
elements outside |C - C64| <= 2 K 2^-24 (|A||B|).  The headline value is the
fastest mode with no such element: the metric is quoted at FP32-level error,
which the one- and two-piece splits (BF16, BF16X3) cannot meet.

The same line carries, under "configs", the metric's other configurations
run in the same call (compact_configs): the 80^3 NT call, the cubic 8..128
sweep, tall-skinny / short-wide sgemm_mp, 100k batched CP2K-sized DGEMMs and
the im2col convolution layers forward + fused backward (batch 32).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-configs]
       python bench.py --workload {conv,mp,sweep,hello,dgemm_batched}
A --workload run prints that configuration's full results as its own line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2516.6 / 3, "bf16x6": 2516.6 / 8, "bf16": 2516.6, "fp64": 78.6}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--count", type=int, default=50000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--modes", default="fp32,bf16x6,bf16x3,bf16",
                    help="precision modes to time; the headline is the fastest one with no sampled result "
                         "outside the FP32 error bound")
    ap.add_argument("--sample", type=int, default=2000,
                    help="GEMMs drawn at random for the error check (plus 8 per K class 1..16)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="batched plans with the fixed tile rule instead of the autotuned tiles")
    ap.add_argument("--no-configs", dest="configs", action="store_false",
                    help="skip the other configurations (Hello 80^3, the cubic sweep, sgemm_mp, the DGEMM batch, "
                         "the conv layers) the default run adds to its JSON line")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "conv", "mp", "sweep", "hello", "dgemm_batched"])
    ap.add_argument("--conv-layers", default="",
                    help="conv workload: comma-separated indices of the layers to run (default all four)")
    ap.add_argument("--mp-shapes", default="",
                    help="mp workload: comma-separated MxNxK shapes of the grid to run (default all)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="issue one device-to-device copy of this many MiB first (untimed), a known-size "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); the batched path "
              f"would shard GEMMs across GPUs with no exchange step", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import uniform_shapes

    dev = torch.device("cuda:0")
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload != "batched":
        emit(OTHER[args.workload](args, dev))
        return
    shapes = uniform_shapes(args.count, 1, 128, seed=0)
    Ms = np.array([s[0] for s in shapes], dtype=np.int64)
    Ns = np.array([s[1] for s in shapes], dtype=np.int64)
    Ks = np.array([s[2] for s in shapes], dtype=np.int64)
    offA = np.concatenate([[0], np.cumsum(Ms * Ks)])
    offB = np.concatenate([[0], np.cumsum(Ks * Ns)])
    offC = np.concatenate([[0], np.cumsum(Ms * Ns)])
    gen = torch.Generator(device=dev).manual_seed(0)
    dA = torch.rand(int(offA[-1]) + 64, device=dev, generator=gen).mul_(2).sub_(1)
    dB = torch.rand(int(offB[-1]) + 64, device=dev, generator=gen).mul_(2).sub_(1)
    dC = torch.zeros(int(offC[-1]) + 64, device=dev)
    # 16-byte aligned starts where the row length allows it
    As = [dA.data_ptr() + 4 * int(o) for o in offA[:-1]]
    Bs = [dB.data_ptr() + 4 * int(o) for o in offB[:-1]]
    Cs = [dC.data_ptr() + 4 * int(o) for o in offC[:-1]]
    flops = float(np.sum(2.0 * Ms * Ns * Ks))
    nbytes = float(np.sum(4.0 * (Ms * Ks + Ks * Ns + Ms * Ns)))
    log(f"workload: {args.count} SGEMMs, M,N,K~U{{1..128}}: {flops / 1e9:.2f} GFLOP, {nbytes / 1e9:.3f} GB")

    checker = SampleChecker(dA, dB, dC, offA, offB, offC, Ms, Ns, Ks, args.sample, dev)
    log(f"error sample: {checker.count} GEMMs ({checker.elements} elements of C), every K <= 16 included")

    allmodes = {"fp32": ls.FP32, "bf16x6": ls.BF16X6, "bf16x3": ls.BF16X3, "bf16": ls.BF16}
    modes = [(m, allmodes[m]) for m in args.modes.split(",")]
    stream = torch.cuda.current_stream()
    results = {}
    for name, prec in modes:
        # tile per shape class from the autotuner (untimed; synthetic operands)
        tuned = ls.autotune_batched(prec, 0, 0, Ms, Ns, Ks, stream) if args.autotune else 0
        plan = ls.BatchedPlan(0, 0, Cs, As, Bs, Ms, Ns, Ks, precision=prec)
        settle(lambda: plan.execute(stream))
        for _ in range(args.warmup):
            plan.execute(stream)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            plan.execute(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        _last_busy[0] = time.perf_counter()
        ms = e0.elapsed_time(e1) / args.steps
        mx, mean, viol = checker.errors()
        info = plan.info()
        plan.close()
        r = {"gflops": flops / (ms * 1e-3) / 1e9, "ms_per_step": ms, "wall_s": wall, "tuned_classes": tuned,
             "max_rel_err": mx, "mean_rel_err": mean, "fp32_tol_violations_in_sample": viol,
             "work_items": info["work_items"],
             "hbm_gbs": nbytes / (ms * 1e-3) / 1e9, **checker.last}
        results[name] = r
        log(f"{name:7s}: {ms:.3f} ms/step  {r['gflops']:.0f} GFLOP/s  {r['hbm_gbs']:.0f} GB/s  "
            f"max_rel {mx:.3g} mean_rel {mean:.3g}  viol {viol}")

    # The headline: the fastest mode whose sampled results all meet the FP32
    # bound |C - C64| <= 2 K 2^-24 (|A||B|).  Modes that miss it (bf16x3,
    # bf16) are reported in "modes" only.
    passing = [k for k, v in results.items() if v["fp32_tol_violations_in_sample"] == 0]
    if not passing:
        log("bench.py: no precision mode met the FP32 error bound on the sample; no headline value")
        sys.exit(1)
    vmode = max(passing, key=lambda k: results[k]["gflops"])
    h = results[vmode]
    log(f"headline mode: {vmode} (fastest of {passing} with 0 violations)")

    # Dominant kernel: the 128 x 128-tile persistent kernel of the headline
    # mode, on the GEMMs the fixed tile rule gives it (max(M, N) > 64; the
    # tuned plans move some of them to 64 x 64 tiles).  A plan holding only
    # those GEMMs, built with the fixed rule, is one launch of that kernel --
    # the launch the rocprofv3 summaries in profiles/ describe; HIP events on
    # its stream give its average duration.
    vprec = allmodes[vmode]
    ls.autotune_reset()
    big = [g for g in range(args.count)
           if ls.query_batched_tile(vprec, int(Ms[g]), int(Ns[g]), int(Ks[g])) == (128, 128)]
    bidx = np.array(big, dtype=np.int64)
    kplan = ls.BatchedPlan(0, 0, [Cs[g] for g in big], [As[g] for g in big], [Bs[g] for g in big],
                           Ms[bidx], Ns[bidx], Ks[bidx], precision=vprec)
    kbytes = float(np.sum(4.0 * (Ms[bidx] * Ks[bidx] + Ks[bidx] * Ns[bidx] + Ms[bidx] * Ns[bidx])))
    kflops = float(np.sum(2.0 * Ms[bidx] * Ns[bidx] * Ks[bidx]))
    settle(lambda: kplan.execute(stream))
    for _ in range(args.warmup):
        kplan.execute(stream)
    k0 = torch.cuda.Event(enable_timing=True)
    k1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    k0.record(stream)
    for _ in range(args.steps):
        kplan.execute(stream)
    k1.record(stream)
    torch.cuda.synchronize()
    _last_busy[0] = time.perf_counter()
    kms = k0.elapsed_time(k1) / args.steps
    kplan.close()
    kgbs = kbytes / (kms * 1e-3) / 1e9
    log(f"dominant kernel ({vmode}, 128x128 bucket, {len(big)} GEMMs): {kms:.3f} ms/launch, {kgbs:.0f} GB/s, "
        f"{kflops / (kms * 1e-3) / 1e12:.1f} TFLOP/s")
    traffic = recorded_traffic(vmode)
    copy_gbs = copy_rate(dev, stream)
    log(f"device copy (1 GiB, torch copy_): {copy_gbs:.0f} GB/s read + write; the dominant kernel moves "
        f"{kgbs / copy_gbs:.3f} of that")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(shapes, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    sample = {"gemms": checker.count, "elements": checker.elements, "k_le_16_classes": checker.small_k_classes}
    configs = None
    if args.configs:
        del dA, dB, dC, checker
        torch.cuda.empty_cache()
        configs = compact_configs(args, dev)

    out = {
        "metric": "batched small-GEMM GFLOP/s (bf16x3 vs FP32 MFMA) at FP32-level error; % roofline",
        "value": round(h["gflops"], 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[vmode],
        "data": "synthetic: U[-1,1) fp32 entries, M,N,K~U{1..128} splitmix64(0)",
        "config": {"workload": "batched SGEMM, 50k GEMMs, M,N,K~U{1..128}, NN, row-major",
                   "count": args.count, "gflop_per_step": round(flops / 1e9, 3),
                   "gb_per_step": round(nbytes / 1e9, 4)},
        "value_mode": vmode,
        "value_rule": "fastest mode with 0 elements outside |C-C64| <= 2K 2^-24 (|A||B|) on the error sample",
        "error_sample": sample,
        "speedup_bf16x3_vs_fp32": (round(results["bf16x3"]["gflops"] / results["fp32"]["gflops"], 3)
                                   if "fp32" in results and "bf16x3" in results else None),
        "modes": {k: {kk: (float(f"{vv:.4g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(kgbs / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": f"k_gemm_persistent<{vmode}, NN, 128x128> (the {len(big)} GEMMs the plan puts in "
                               f"128x128 tiles)",
                     "kernel_ms": round(kms, 4), "kernel_algorithmic_bytes": kbytes,
                     "kernel_tflops": round(kflops / (kms * 1e-3) / 1e12, 2),
                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(kgbs / copy_gbs, 3)},
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


DTYPES = {"fp32": "fp32 (fp32 MFMA, fp32 accumulate)",
          "bf16x6": "bf16x6 (fp32 in/out, three-piece bf16 split on bf16 MFMA, fp32 accumulate)",
          "bf16x3": "bf16x3 (fp32 in/out, bf16 MFMA, fp32 accumulate)",
          "bf16": "bf16 (fp32 in/out, bf16 MFMA, fp32 accumulate)"}


class SampleChecker:
    """Componentwise error of a sample of the batch's GEMMs against an FP64
    product: the sampled A, B (and, per mode, C) are gathered into zero-padded
    128 x 128 float64 blocks and multiplied with one batched product on the
    GPU.  The sample is `count` GEMMs drawn at random plus, for every K from 1
    to 16, eight GEMMs with that K (the bound 2 K u is tightest there)."""

    def __init__(self, dA, dB, dC, offA, offB, offC, Ms, Ns, Ks, count, dev):
        import numpy as np
        import torch
        rng = np.random.default_rng(1)
        n = len(Ms)
        pick = list(rng.choice(n, size=min(count, n), replace=False))
        have = set(int(g) for g in pick)
        self.small_k_classes = 0
        for k in range(1, 17):
            extra = [int(g) for g in np.nonzero(Ks == k)[0] if int(g) not in have][:8]
            if extra or any(int(Ks[g]) == k for g in pick):
                self.small_k_classes += 1
            pick += extra
            have.update(extra)
        pick = np.array(sorted(int(g) for g in pick), dtype=np.int64)
        self.count = len(pick)
        T = 128

        def index(off, rows, cols):
            src, dst = [], []
            for s, g in enumerate(pick):
                r, c = int(rows[g]), int(cols[g])
                i = np.arange(r * c, dtype=np.int64)
                src.append(int(off[g]) + i)
                dst.append(s * T * T + (i // c) * T + i % c)
            return (torch.from_numpy(np.concatenate(src)).to(dev), torch.from_numpy(np.concatenate(dst)).to(dev))

        sa, da = index(offA, Ms, Ks)
        sb, db = index(offB, Ks, Ns)
        self.sc, self.dc = index(offC, Ms, Ns)
        S = self.count
        A = torch.zeros(S * T * T, dtype=torch.float64, device=dev)
        B = torch.zeros(S * T * T, dtype=torch.float64, device=dev)
        A[da] = dA[sa].double()
        B[db] = dB[sb].double()
        A = A.view(S, T, T)
        B = B.view(S, T, T)
        self.ref = torch.bmm(A, B)
        self.den = torch.bmm(A.abs(), B.abs())
        self.bound = self.den * (2.0 * 2.0 ** -24) * torch.from_numpy(Ks[pick].astype(np.float64)).to(dev).view(S, 1, 1)
        del A, B
        self.dCflat = dC
        self.elements = int(self.sc.numel())
        self.Ks = torch.from_numpy(Ks[pick]).to(dev)
        self.last = {}

    def errors(self):
        import torch
        S = self.count
        C = torch.zeros(S * 128 * 128, dtype=torch.float64, device=self.ref.device)
        C[self.dc] = self.dCflat[self.sc].double()
        e = (C.view(S, 128, 128) - self.ref).abs()
        mask = self.den > 0
        rel = e[mask] / self.den[mask]
        over = e > self.bound
        viol = int(over.sum())
        # worst |C - C64| / (2 K u |A||B|) and the K of the GEMMs holding violations
        ratio = (e[mask] / self.bound[mask]).max()
        per = over.view(S, -1).sum(1)
        ks = sorted(set(self.Ks[per > 0].tolist()))
        self.last = {"max_ratio_to_bound": float(ratio),
                     "violating_k_range": [ks[0], ks[-1]] if ks else None,
                     "violations_k_le_4": {str(k): int(per[self.Ks == k].sum()) for k in ks if k <= 4}}
        return float(rel.max()), float(rel.mean()), viol


def recorded_traffic(mode):
    """HBM bytes per launch of the dominant kernel of `mode` as measured by
    rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy) and
    recorded in profiles/; None when no such record exists."""
    path = os.path.join(ROOT, "profiles", f"batched_sgemm_{mode}_128_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(shapes, seconds):
    """The oracle's OpenMP triple loop on a bounded sample of the same batch."""
    import numpy as np

    import oracle

    rng = np.random.default_rng(2)
    n = 0
    sel = []
    tot = 0.0
    order = rng.permutation(len(shapes))
    # ~5 GFLOP sample, or fewer GEMMs if the batch is small
    for g in order:
        m, nn, k = shapes[g]
        sel.append(shapes[g])
        tot += 2.0 * m * nn * k
        n += 1
        if tot > 5e9:
            break
    As = [(rng.random((m, k), dtype=np.float32) * 2 - 1) for m, nn, k in sel]
    Bs = [(rng.random((k, nn), dtype=np.float32) * 2 - 1) for m, nn, k in sel]
    Cs = [np.zeros((m, nn), np.float32) for m, nn, k in sel]
    Ms, Ns, Ks = zip(*sel)
    oracle.cpu_gemm_batched(0, 0, Cs, As, Bs, Ms, Ns, Ks)  # warm-up
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.cpu_gemm_batched(0, 0, Cs, As, Bs, Ms, Ns, Ks)
        reps += 1
        el = time.perf_counter() - t0
        if el > seconds or reps >= 50:
            break
    gflops = tot * reps / el / 1e9
    return {"value": round(gflops, 2), "unit": "GFLOP/s", "cores": oracle.cpu_threads(), "kind": "port",
            "sample": f"{n} GEMMs of the batch ({tot / 1e9:.2f} GFLOP) x {reps} reps, FP32 OpenMP triple loop "
                      f"(i-k-j), gcc -O3 -fopenmp -march=x86-64-v3"}


# ---------------------------------------------------------------------------
# Other workloads (one JSON line each).
# ---------------------------------------------------------------------------
_last_busy = [0.0]


def settle(fn):
    """Run fn back to back, untimed, until the GPU's clocks have ramped up.
    After an idle gap the first tens of milliseconds of load run at a lower
    clock (measured: the first 20 launches of the batched FP32 128x128 kernel
    average 0.646 ms, every later group 0.563 ms; again after 0.5 s idle), so
    each timed region is preceded by 0.2 s of its own work, or 0.03 s when the
    previous one ended less than 0.1 s ago.  The timed steps and the warmup
    steps are unchanged."""
    import torch
    t0 = time.perf_counter()
    span = 0.03 if t0 - _last_busy[0] < 0.1 else 0.2
    while time.perf_counter() - t0 < span:
        for _ in range(4):
            fn()
        torch.cuda.synchronize()


def time_ms(fn, steps, warmup, stream):
    import torch
    settle(fn)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    _last_busy[0] = time.perf_counter()
    return e0.elapsed_time(e1) / steps


def copy_rate(dev, stream):
    """GB/s (bytes read + bytes written) of a plain 1 GiB device-to-device
    copy on this GPU: what a kernel with no arithmetic reaches, beside the
    8 TB/s peak the roofline fraction is taken against."""
    import torch
    n = 256 << 20
    src = torch.ones(n, device=dev)
    dst = torch.empty_like(src)
    ms = time_ms(lambda: dst.copy_(src), 10, 2, stream)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * 4 * n / (ms * 1e-3) / 1e9


def roof_tflops(mode, flops, nbytes):
    """min(MFMA peak of the mode, arithmetic intensity x HBM peak), TFLOP/s."""
    return min(PEAK_TFLOPS[mode], flops / nbytes * PEAK_HBM_GBS / 1e3)


def rel_err(got, ref, den):
    e = (got.double() - ref).abs()
    r = e / den.clamp_min(1e-300)
    return float(r.max()), float(r.mean()), e, den


def emit(obj):
    print(json.dumps(obj), flush=True)


def bench_conv(args, dev):
    """im2col convolution layers, batch 32: forward C_i = W X_i (one strided
    batched launch, W shared) and the fused backward (dW summed over the
    batch, dX per image) in one launch.  W ~ N(0, 2/K), X and dC ~ N(0, 1)."""
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import CONV_SHAPES
    stream = torch.cuda.current_stream()
    batch = 32
    layers, tot_f, tot_ms = [], {}, {}
    picked = [int(x) for x in args.conv_layers.split(",")] if args.conv_layers else range(len(CONV_SHAPES))
    for (Cout, HW, Kc) in [CONV_SHAPES[i] for i in picked]:
        g = torch.Generator(device=dev).manual_seed(Cout)
        W = torch.randn(Cout, Kc, generator=g, device=dev) * (2.0 / Kc) ** 0.5
        X = torch.randn(batch, Kc, HW, generator=g, device=dev)
        dY = torch.randn(batch, Cout, HW, generator=g, device=dev)
        Y = torch.empty(batch, Cout, HW, device=dev)
        dW = torch.empty_like(W)
        dX = torch.empty_like(X)
        ff = 2.0 * Cout * HW * Kc * batch
        fb = 2 * ff
        bf = 4.0 * (Cout * Kc + batch * Kc * HW + batch * Cout * HW)
        bb = 4.0 * (batch * Cout * HW + 2 * Cout * Kc + 2 * batch * Kc * HW)
        row = {"M": Cout, "N": HW, "K": Kc, "batch": batch}
        for name, prec in (("fp32", ls.FP32), ("bf16x3", ls.BF16X3)):
            fms = time_ms(lambda: ls.sgemm_strided_batched(0, 0, Y, W, X, Cout, HW, Kc, batch, 0, Kc * HW,
                                                           Cout * HW, prec, stream),
                          args.steps, args.warmup, stream)
            bms = time_ms(lambda: ls.sgemm_backward_strided(0, 0, dY, W, X, dW, dX, Cout, HW, Kc, batch, 0,
                                                            Kc * HW, Cout * HW, prec, stream),
                          args.steps, args.warmup, stream)
            # errors: forward and dX on image 0, dW in full
            W64, X0, dY64 = W.double(), X[0].double(), dY.double()
            fe = rel_err(Y[0], W64 @ X0, W64.abs() @ X0.abs())
            gW = sum(dY64[i] @ X[i].double().T for i in range(batch))
            bW = sum(dY64[i].abs() @ X[i].double().abs().T for i in range(batch))
            we = rel_err(dW, gW, bW)
            xe = rel_err(dX[0], W64.T @ dY64[0], W64.abs().T @ dY64[0].abs())
            row[name] = {
                "fwd_ms": round(fms, 4), "bwd_ms": round(bms, 4),
                "fwd_tflops": round(ff / fms / 1e9, 2), "bwd_tflops": round(fb / bms / 1e9, 2),
                "fwd_roof_frac": round(ff / fms / 1e9 / roof_tflops(name, ff, bf), 3),
                "bwd_roof_frac": round(fb / bms / 1e9 / roof_tflops(name, fb, bb), 3),
                "fwd_max_rel": float(f"{fe[0]:.3g}"), "dW_max_rel": float(f"{we[0]:.3g}"),
                "dX_max_rel": float(f"{xe[0]:.3g}"),
            }
            # implicit GEMM: the im2col matrix gathered by the operand loads from
            # NCHW images (3 x 3, stride 1, padding 1), against torch's unfold
            # followed by the same strided GEMM
            Cin, side = Kc // 9, int(round(HW ** 0.5))
            Ximg = torch.randn(batch, Cin, side, side, generator=g, device=dev)
            ims = time_ms(lambda: ls.sconv2d_forward(Y, W, Ximg, batch, Cin, side, side, Cout, 3, 3, (1, 1),
                                                     (1, 1), (1, 1), prec, stream),
                          args.steps, args.warmup, stream)

            def explicit():
                cols = torch.nn.functional.unfold(Ximg, 3, padding=1).contiguous()
                ls.sgemm_strided_batched(0, 0, Y, W, cols, Cout, HW, Kc, batch, 0, Kc * HW, Cout * HW, prec,
                                         stream)
            ems = time_ms(explicit, args.steps, args.warmup, stream)
            ls.sconv2d_forward(Y, W, Ximg, batch, Cin, side, side, Cout, 3, 3, (1, 1), (1, 1), (1, 1), prec,
                               stream)
            torch.cuda.synchronize()
            c64 = torch.nn.functional.unfold(Ximg[:1].double(), 3, padding=1)[0]
            ie = rel_err(Y[0], W64 @ c64, W64.abs() @ c64.abs())
            bi = 4.0 * (Cout * Kc + batch * Cin * HW + batch * Cout * HW)
            row[name].update({
                "implicit_fwd_ms": round(ims, 4), "implicit_fwd_tflops": round(ff / ims / 1e9, 2),
                "implicit_fwd_roof_frac": round(ff / ims / 1e9 / roof_tflops(name, ff, bi), 3),
                "explicit_unfold_plus_fwd_ms": round(ems, 4), "implicit_max_rel": float(f"{ie[0]:.3g}"),
            })
            log(f"conv {Cout}x{HW}x{Kc} {name}: implicit fwd {ims:.3f} ms ({ff / ims / 1e9:.1f} TF/s), "
                f"unfold + fwd {ems:.3f} ms, max rel {ie[0]:.2g}")
            # implicit backward (dX images and dW in one launch) against unfold + the
            # fused stored-im2col backward + fold (col2im) of the column gradient
            dXimg = torch.empty_like(Ximg)

            def implicit_bwd():
                ls.sconv2d_backward(dXimg, dW, dY, W, Ximg, batch, Cin, side, side, Cout, 3, 3, (1, 1), (1, 1),
                                    (1, 1), prec, stream)
            ibms = time_ms(implicit_bwd, args.steps, args.warmup, stream)

            def explicit_bwd():
                cols = torch.nn.functional.unfold(Ximg, 3, padding=1).contiguous()
                ls.sgemm_backward_strided(0, 0, dY, W, cols, dW, dX, Cout, HW, Kc, batch, 0, Kc * HW, Cout * HW,
                                          prec, stream)
                torch.nn.functional.fold(dX, (side, side), 3, padding=1)
            ebms = time_ms(explicit_bwd, args.steps, args.warmup, stream)
            implicit_bwd()
            torch.cuda.synchronize()
            dc64 = W64.T @ dY64[0]
            rx = torch.nn.functional.fold(dc64[None], (side, side), 3, padding=1)[0]
            ax = torch.nn.functional.fold((W64.abs().T @ dY64[0].abs())[None], (side, side), 3, padding=1)[0]
            dxe = rel_err(dXimg[0], rx, ax)
            gW64 = torch.zeros_like(W64)
            aW64 = torch.zeros_like(W64)
            for i0 in range(0, batch, 8):
                c64 = torch.nn.functional.unfold(Ximg[i0:i0 + 8].double(), 3, padding=1)
                ct = c64.transpose(1, 2).contiguous()
                gW64 += torch.matmul(dY64[i0:i0 + 8], ct).sum(0)
                aW64 += torch.matmul(dY64[i0:i0 + 8].abs(), ct.abs()).sum(0)
            dwe = rel_err(dW, gW64, aW64)
            row[name].update({
                "implicit_bwd_ms": round(ibms, 4), "implicit_bwd_tflops": round(fb / ibms / 1e9, 2),
                "explicit_unfold_bwd_fold_ms": round(ebms, 4),
                "implicit_dX_max_rel": float(f"{dxe[0]:.3g}"), "implicit_dW_max_rel": float(f"{dwe[0]:.3g}"),
            })
            log(f"conv {Cout}x{HW}x{Kc} {name}: implicit bwd {ibms:.3f} ms ({fb / ibms / 1e9:.1f} TF/s), "
                f"unfold + bwd + fold {ebms:.3f} ms, max rel dX {dxe[0]:.2g} dW {dwe[0]:.2g}")
            del Ximg, dXimg
            tot_f[name] = tot_f.get(name, 0.0) + ff + fb
            tot_ms[name] = tot_ms.get(name, 0.0) + fms + bms
            log(f"conv {Cout}x{HW}x{Kc} {name}: fwd {fms:.3f} ms ({ff / fms / 1e9:.1f} TF/s), "
                f"bwd {bms:.3f} ms ({fb / bms / 1e9:.1f} TF/s), max rel fwd {fe[0]:.2g} dW {we[0]:.2g} "
                f"dX {xe[0]:.2g}")
        layers.append(row)
        del W, X, dY, Y, dW, dX
        torch.cuda.empty_cache()
    return ({"metric": "im2col conv GEMM forward + fused backward GFLOP/s, batch 32 (four layers)",
          "value": round(tot_f["bf16x3"] / tot_ms["bf16x3"] / 1e6, 1), "unit": "GFLOP/s",
          "fp32_gflops": round(tot_f["fp32"] / tot_ms["fp32"] / 1e6, 1), "n_gpus": 1,
          "dtype": "bf16x3 (value) and fp32", "data": "synthetic: W~N(0,2/K), X,dC~N(0,1)",
          "config": {"workload": "conv layers (M=Cout,N=HW,K=Cin*9) x batch 32"}, "layers": layers})


def bench_mp(args, dev):
    """LibShalom_sgemm_mp shapes: tall-skinny M x N x K and short-wide
    transposes, NN and NT, FP32 vs BF16X3, per-shape GFLOP/s, fraction of
    min(MFMA peak, AI x HBM peak), and max / mean relative error on 64
    sampled rows against FP64."""
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import TALL_SKINNY
    stream = torch.cuda.current_stream()
    g = torch.Generator(device=dev).manual_seed(0)
    big = torch.rand(65536 * 1024 + 64, generator=g, device=dev).mul_(2).sub_(1)
    bigB = torch.rand(65536 * 1024 + 64, generator=g, device=dev).mul_(2).sub_(1)
    Cbuf = torch.empty(65536 * 1024, device=dev)
    shapes = [(M, N, K, "tall") for (M, N, K) in TALL_SKINNY]
    shapes += [(N, M, K, "wide") for (M, N, K) in TALL_SKINNY]  # short-wide: M small, N large
    if getattr(args, "mp_shapes", ""):
        keep = {tuple(int(v) for v in x.split("x")) for x in args.mp_shapes.split(",")}
        shapes = [s for s in shapes if s[:3] in keep]
    rows = []
    rng = torch.Generator().manual_seed(1)
    for (M, N, K, kind) in shapes:
        A = big[:M * K].view(M, K)
        for tb in (0, 1):
            B = bigB[:K * N].view(N, K) if tb else bigB[:K * N].view(K, N)
            C = Cbuf[:M * N].view(M, N)
            opB = B.double().T if tb else B.double()
            sel = torch.randint(0, M, (min(64, M),), generator=rng).to(dev)
            row = {"M": M, "N": N, "K": K, "tb": tb, "kind": kind}
            flops = 2.0 * M * N * K
            nbytes = 4.0 * (M * K + K * N + M * N)
            for name, prec in (("fp32", ls.FP32), ("bf16x3", ls.BF16X3)):
                ms = time_ms(lambda: ls.sgemm_mp_ex(0, tb, C, A, B, M, N, K, prec, stream), args.steps,
                             args.warmup, stream)
                a = A[sel].double()
                e = rel_err(C[sel], a @ opB, a.abs() @ opB.abs())
                tf = flops / ms / 1e9
                row[name] = {"ms": round(ms, 5), "tflops": round(tf, 2),
                             "roof_frac": round(tf / roof_tflops(name, flops, nbytes), 3),
                             "max_rel": float(f"{e[0]:.3g}"), "mean_rel": float(f"{e[1]:.3g}")}
            rows.append(row)
            log(f"mp {kind} {M}x{N}x{K} tb={tb}: fp32 {row['fp32']['tflops']} TF/s "
                f"({row['fp32']['roof_frac']}), bf16x3 {row['bf16x3']['tflops']} TF/s "
                f"({row['bf16x3']['roof_frac']}), rel err fp32 {row['fp32']['max_rel']} "
                f"bf16x3 {row['bf16x3']['max_rel']}")
    # the reference ABI on the 4096 x 16 x 1024 NN shape: device operands
    # through the synchronous call, and host operands (staged through pinned
    # bounce buffers, the 16 MiB A dominating)
    import numpy as np
    M, N, K = 4096, 16, 1024
    A = big[:M * K].view(M, K)
    B = bigB[:K * N].view(K, N)
    C = Cbuf[:M * N].view(M, N)
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    Ch = np.empty((M, N), np.float32)
    ref_api = {}
    for name, (a, b, c) in (("device_sync_us", (A, B, C)), ("host_pointer_us", (Ah, Bh, Ch))):
        ls.sgemm_mp(0, 0, c, a, b, M, N, K)
        t0 = time.perf_counter()
        for _ in range(20):
            ls.sgemm_mp(0, 0, c, a, b, M, N, K)
        ref_api[name] = round((time.perf_counter() - t0) / 20 * 1e6, 1)
    assert np.array_equal(Ch, C.cpu().numpy())
    log(f"mp reference API 4096x16x1024: {ref_api}")
    tf = sum(2.0 * r["M"] * r["N"] * r["K"] for r in rows)
    t32 = sum(r["fp32"]["ms"] for r in rows)
    t3 = sum(r["bf16x3"]["ms"] for r in rows)
    return ({"metric": "sgemm_mp tall-skinny / short-wide GFLOP/s (all shapes, NN+NT)",
          "value": round(tf / t3 / 1e6, 1), "unit": "GFLOP/s", "fp32_gflops": round(tf / t32 / 1e6, 1),
          "n_gpus": 1, "dtype": "bf16x3 (value) and fp32", "data": "synthetic: U[-1,1)",
          "config": {"workload": "M in {4096,16384,65536} x N in {16,32,64} x K in {64,256,1024} and "
                                 "the transposed short-wide shapes"}, "reference_api_4096x16x1024": ref_api,
          "shapes": rows})


def bench_sweep(args, dev):
    """Cubic n = 8..128 step 8, the four NoTrans/Trans combinations, FP32:
    one LibShalom_sgemm_ex call per size, 100 timed calls after 5 warm-ups."""
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import SMALL_SWEEP
    stream = torch.cuda.current_stream()
    rows = []
    g = torch.Generator(device=dev).manual_seed(0)
    for (n, _, _) in SMALL_SWEEP:
        A = torch.rand(n, n, generator=g, device=dev).mul_(2).sub_(1)
        B = torch.rand(n, n, generator=g, device=dev).mul_(2).sub_(1)
        C = torch.empty(n, n, device=dev)
        row = {"n": n}
        for ta in (0, 1):
            for tb in (0, 1):
                C.fill_(float("nan"))
                ms = time_ms(lambda: ls.sgemm_ex(ta, tb, C, A, B, n, n, n, ls.FP32, stream), 100, 5, stream)
                a = A.double().T if ta else A.double()
                b = B.double().T if tb else B.double()
                _, _, e, den = rel_err(C, a @ b, a.abs() @ b.abs())
                viol = int((e > 2 * n * 2.0 ** -24 * den).sum()) + int(torch.isnan(C).sum())
                row[f"{'NT'[ta]}{'NT'[tb]}"] = {"us": round(ms * 1e3, 2),
                                                "gflops": round(2.0 * n ** 3 / ms / 1e6, 2), "viol": viol}
        rows.append(row)
        log(f"sweep n={n}: " + " ".join(f"{k} {v['us']}us" for k, v in row.items() if k != "n"))
    cpu = None
    if not args.no_cpu:
        import numpy as np

        import oracle
        cpu = []
        for (n, _, _) in SMALL_SWEEP:
            A = np.random.default_rng(n).random((n, n), dtype=np.float32)
            t0 = time.perf_counter()
            for _ in range(100):
                oracle.cpu_sgemm(0, 0, A, A, n, n, n)
            el = (time.perf_counter() - t0) / 100
            cpu.append({"n": n, "us": round(el * 1e6, 2), "gflops": round(2.0 * n ** 3 / el / 1e9, 2)})
        cpu = {"kind": "port", "cores": oracle.cpu_threads(), "sizes": cpu}
    return ({"metric": "small SGEMM sweep: per-call time, n=8..128 step 8, FP32", "value": rows[-1]["NN"]["gflops"],
          "unit": "GFLOP/s at n=128 NN", "n_gpus": 1, "dtype": "fp32", "data": "synthetic: U[-1,1)",
          "config": {"workload": "cubic n=8..128 step 8 x {NN,NT,TN,TT}, 5 warm-ups + 100 calls"},
          "sizes": rows, "cpu_baseline": cpu})


def bench_hello(args, dev):
    """The 80^3 NoTrans/Trans call with the reference harness's inputs."""
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import hello_inputs
    A, B = hello_inputs(80, 80, 80)
    At = torch.from_numpy(A).to(dev)
    Bt = torch.from_numpy(B).to(dev)
    C = torch.full((80, 80), float("nan"), device=dev)
    stream = torch.cuda.current_stream()
    ms = time_ms(lambda: ls.sgemm_ex(0, 1, C, At, Bt, 80, 80, 80, ls.FP32, stream), 100, 5, stream)
    t0 = time.perf_counter()
    for _ in range(100):
        ls.sgemm(0, 1, C, At, Bt, 80, 80, 80)
    sync_us = (time.perf_counter() - t0) / 100 * 1e6
    # the reference's own ABI: host buffers (staged through the device)
    import numpy as np
    Ch = np.empty((80, 80), np.float32)
    ls.sgemm(0, 1, Ch, A, B, 80, 80, 80)
    t0 = time.perf_counter()
    for _ in range(100):
        ls.sgemm(0, 1, Ch, A, B, 80, 80, 80)
    host_us = (time.perf_counter() - t0) / 100 * 1e6
    a = At.double().view(80, 80)
    b = Bt.double().view(80, 80).T
    mx, mean, e, den = rel_err(C, a @ b, a.abs() @ b.abs())
    viol = int((e > 2 * 80 * 2.0 ** -24 * den).sum())
    return ({"metric": "Hello 80x80x80 NT sgemm", "value": round(ms * 1e3, 2), "unit": "us per call (stream)",
          "higher_is_better": False, "sync_api_us": round(sync_us, 2), "host_pointer_api_us": round(host_us, 2),
          "gflops": round(1.024e6 / ms / 1e6, 2),
          "violations": viol, "max_rel": float(f"{mx:.3g}"), "n_gpus": 1, "dtype": "fp32",
          "data": "drand48 default-state inputs (reference harness recipe)",
          "config": {"workload": "LibShalom_sgemm(NoTrans, Trans) M=N=K=80"}})


def bench_dgemm_batched(args, dev):
    """100,000 DGEMMs with M, N, K drawn from the small-block dims
    {4,5,9,13,16,22,23,26,32}, FP64 on FP64 MFMA, one batched plan."""
    import numpy as np
    import torch

    import libshalom_amd as ls
    from libshalom_amd.workloads import cp2k_shapes
    count = 100000
    shapes = cp2k_shapes(count, seed=0)
    Ms = np.array([s[0] for s in shapes], dtype=np.int64)
    Ns = np.array([s[1] for s in shapes], dtype=np.int64)
    Ks = np.array([s[2] for s in shapes], dtype=np.int64)
    offA = np.concatenate([[0], np.cumsum(Ms * Ks)])
    offB = np.concatenate([[0], np.cumsum(Ks * Ns)])
    offC = np.concatenate([[0], np.cumsum(Ms * Ns)])
    g = torch.Generator(device=dev).manual_seed(0)
    dA = torch.rand(int(offA[-1]) + 8, generator=g, device=dev, dtype=torch.float64).mul_(2).sub_(1)
    dB = torch.rand(int(offB[-1]) + 8, generator=g, device=dev, dtype=torch.float64).mul_(2).sub_(1)
    dC = torch.zeros(int(offC[-1]) + 8, device=dev, dtype=torch.float64)
    As = [dA.data_ptr() + 8 * int(o) for o in offA[:-1]]
    Bs = [dB.data_ptr() + 8 * int(o) for o in offB[:-1]]
    Cs = [dC.data_ptr() + 8 * int(o) for o in offC[:-1]]
    flops = float(np.sum(2.0 * Ms * Ns * Ks))
    nbytes = float(np.sum(8.0 * (Ms * Ks + Ks * Ns + Ms * Ns)))
    stream = torch.cuda.current_stream()
    plan = ls.BatchedPlan(0, 0, Cs, As, Bs, Ms, Ns, Ks, fp64=True)
    ms = time_ms(lambda: plan.execute(stream), args.steps, args.warmup, stream)
    plan.close()
    mx = 0.0
    for i in np.random.default_rng(1).choice(count, 300, replace=False):
        m, n, k = int(Ms[i]), int(Ns[i]), int(Ks[i])
        a = dA[offA[i]:offA[i] + m * k].view(m, k)
        b = dB[offB[i]:offB[i] + k * n].view(k, n)
        c = dC[offC[i]:offC[i] + m * n].view(m, n)
        mx = max(mx, rel_err(c, a @ b, a.abs() @ b.abs())[0])
    gbs = nbytes / ms / 1e6
    cpu = None
    if not args.no_cpu:
        import oracle
        sel = list(range(2000))
        Ah = [np.random.default_rng(i).random((int(Ms[i]), int(Ks[i]))) for i in sel]
        Bh = [np.random.default_rng(i + 1).random((int(Ks[i]), int(Ns[i]))) for i in sel]
        Ch = [np.zeros((int(Ms[i]), int(Ns[i]))) for i in sel]
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 3.0:
            oracle.cpu_gemm_batched(0, 0, Ch, Ah, Bh, Ms[sel], Ns[sel], Ks[sel])
            reps += 1
        el = time.perf_counter() - t0
        sf = float(np.sum(2.0 * Ms[sel] * Ns[sel] * Ks[sel]))
        cpu = {"value": round(sf * reps / el / 1e9, 2), "unit": "GFLOP/s", "cores": oracle.cpu_threads(),
               "kind": "port", "sample": f"first 2000 GEMMs x {reps} reps, FP64 OpenMP triple loop"}
    return ({"metric": "batched DGEMM GFLOP/s, 100k GEMMs with CP2K block dims", "value": round(flops / ms / 1e6, 1),
          "unit": "GFLOP/s", "ms_per_step": round(ms, 4), "hbm_gbs": round(gbs, 1),
          "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": round(gbs / PEAK_HBM_GBS, 4)},
          "max_rel_err": float(f"{mx:.3g}"), "n_gpus": 1, "dtype": "f64", "data": "synthetic: U[-1,1)",
          "config": {"workload": "100k DGEMMs, M,N,K from {4,5,9,13,16,22,23,26,32}, NN",
                     "gflop_per_step": round(flops / 1e9, 4), "gb_per_step": round(nbytes / 1e9, 4)},
          "cpu_baseline": cpu})


def r3(x):
    return float(f"{x:.3g}")


def compact_configs(args, dev):
    """Every configuration of the metric's benchmark set, run in the default
    bench.py call and summarised compactly for its one JSON line: cfg1 the
    80^3 NT call, cfg2 the cubic sweep (microseconds per call, four trans
    combinations), cfg3 the tall-skinny / short-wide sgemm_mp grid (per shape:
    M, N, K, transb, then for FP32 and BF16X3 microseconds, fraction of
    min(MFMA peak, AI x HBM peak) and max relative error), cfg4a the 100k
    CP2K-sized DGEMM batch, cfg5 the convolution layers (per layer and mode:
    stored-im2col forward / fused backward and implicit forward / backward
    times, roof fractions, max relative errors)."""
    out = {}
    t0 = time.perf_counter()
    h = bench_hello(args, dev)
    out["cfg1_hello_80_nt"] = {k: h[k] for k in ("value", "unit", "sync_api_us", "host_pointer_api_us", "gflops",
                                                  "violations", "max_rel")}
    log(f"cfg1: {out['cfg1_hello_80_nt']}")
    sw = bench_sweep(args, dev)
    rows = sw["sizes"]
    out["cfg2_sweep"] = {"n": [r["n"] for r in rows],
                         **{f"{t}_us": [r[t]["us"] for r in rows] for t in ("NN", "NT", "TN", "TT")},
                         "violations": sum(r[t]["viol"] for r in rows for t in ("NN", "NT", "TN", "TT")),
                         "cpu_us_NN": [c["us"] for c in sw["cpu_baseline"]["sizes"]] if sw["cpu_baseline"] else None,
                         "cpu_cores": sw["cpu_baseline"]["cores"] if sw["cpu_baseline"] else None}
    log(f"cfg2: n=128 NN {rows[-1]['NN']['us']} us, violations {out['cfg2_sweep']['violations']}")
    mp = bench_mp(args, dev)
    out["cfg3_mp"] = {"gflops_bf16x3": mp["value"], "gflops_fp32": mp["fp32_gflops"],
                      "reference_api_4096x16x1024": mp["reference_api_4096x16x1024"],
                      "columns": "M,N,K,tb,fp32_us,fp32_roof_frac,fp32_max_rel,bf16x3_us,bf16x3_roof_frac,"
                                 "bf16x3_max_rel",
                      "max_rel_fp32": max(r["fp32"]["max_rel"] for r in mp["shapes"]),
                      "mean_rel_fp32": r3(sum(r["fp32"]["mean_rel"] for r in mp["shapes"]) / len(mp["shapes"])),
                      "max_rel_bf16x3": max(r["bf16x3"]["max_rel"] for r in mp["shapes"]),
                      "mean_rel_bf16x3": r3(sum(r["bf16x3"]["mean_rel"] for r in mp["shapes"]) / len(mp["shapes"])),
                      "rows": [[r["M"], r["N"], r["K"], r["tb"], round(r["fp32"]["ms"] * 1e3, 1),
                                r["fp32"]["roof_frac"], r3(r["fp32"]["max_rel"]), round(r["bf16x3"]["ms"] * 1e3, 1),
                                r["bf16x3"]["roof_frac"], r3(r["bf16x3"]["max_rel"])] for r in mp["shapes"]]}
    log(f"cfg3: bf16x3 {mp['value']} GFLOP/s, fp32 {mp['fp32_gflops']} GFLOP/s")
    dg = bench_dgemm_batched(args, dev)
    out["cfg4a_dgemm_batch"] = {"gflops": dg["value"], "ms_per_step": dg["ms_per_step"], "hbm_gbs": dg["hbm_gbs"],
                                "hbm_frac": dg["roofline"]["frac"], "max_rel_err": dg["max_rel_err"],
                                "cpu_baseline": dg["cpu_baseline"]}
    log(f"cfg4a: {out['cfg4a_dgemm_batch']}")
    cv = bench_conv(args, dev)
    keys = ("fwd_ms", "fwd_roof_frac", "bwd_ms", "bwd_roof_frac", "implicit_fwd_ms", "implicit_fwd_roof_frac",
            "implicit_bwd_ms", "fwd_max_rel", "dW_max_rel", "dX_max_rel", "implicit_dX_max_rel",
            "implicit_dW_max_rel")
    out["cfg5_conv"] = {"gflops_bf16x3": cv["value"], "gflops_fp32": cv["fp32_gflops"], "columns": ",".join(keys),
                        "layers": {f"{l['M']}x{l['N']}x{l['K']}": {m: [l[m][k] for k in keys]
                                                                  for m in ("fp32", "bf16x3")}
                                   for l in cv["layers"]}}
    log(f"cfg5: bf16x3 {cv['value']} GFLOP/s, fp32 {cv['fp32_gflops']} GFLOP/s")
    out["wall_s"] = round(time.perf_counter() - t0, 1)
    return out


OTHER = {"conv": bench_conv, "mp": bench_mp, "sweep": bench_sweep, "hello": bench_hello,
         "dgemm_batched": bench_dgemm_batched}


if __name__ == "__main__":
    main()
