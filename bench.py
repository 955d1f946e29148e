"""Benchmark of the batched small-GEMM path on one MI355X.

Workload (the metric's configuration): 50,000 independent SGEMMs, C = A*B
(NoTrans/NoTrans, row-major), with M, N, K each uniform on {1..128}
(splitmix64 seed 0), entries uniform in [-1, 1), all resident in HBM.  One
step = one replay of the batched plan (a hipGraph of the size-bucketed tile
kernels) over the whole batch.  The headline mode is BF16X3 (split-BF16 on
BF16 MFMA, FP32 accumulation); FP32 MFMA, BF16X6 and BF16 are timed the same
way and reported beside it, each with its max / mean componentwise relative
error |C - C64| / (|A||B|) against an FP64 product of a sample of the GEMMs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2516.6 / 3, "bf16x6": 2516.6 / 6, "bf16": 2516.6, "fp64": 78.6}


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
    ap.add_argument("--modes", default="bf16x3,fp32,bf16x6,bf16",
                    help="precision modes to time besides the headline bf16x3")
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

    # FP64 reference for a sample of GEMMs (torch on the GPU, fp64)
    rng = np.random.default_rng(1)
    sample = rng.choice(args.count, size=min(300, args.count), replace=False)

    def errors():
        mx, sm, cnt, viol = 0.0, 0.0, 0, 0
        for g in sample:
            m, n, k = int(Ms[g]), int(Ns[g]), int(Ks[g])
            a = dA[offA[g]:offA[g] + m * k].view(m, k).double()
            b = dB[offB[g]:offB[g] + k * n].view(k, n).double()
            c = dC[offC[g]:offC[g] + m * n].view(m, n).double()
            ref = a @ b
            den = a.abs() @ b.abs()
            e = (c - ref).abs()
            rel = (e / den).flatten()
            mx = max(mx, float(rel.max()))
            sm += float(rel.sum())
            cnt += rel.numel()
            viol += int((e > 2 * k * 2.0 ** -24 * den).sum())
        return mx, sm / cnt, viol

    allmodes = {"bf16x3": ls.BF16X3, "fp32": ls.FP32, "bf16x6": ls.BF16X6, "bf16": ls.BF16}
    modes = [(m, allmodes[m]) for m in args.modes.split(",")]
    if "bf16x3" not in args.modes.split(","):
        modes = [("bf16x3", ls.BF16X3)] + modes
    stream = torch.cuda.current_stream()
    results = {}
    headline = None
    for name, prec in modes:
        plan = ls.BatchedPlan(0, 0, Cs, As, Bs, Ms, Ns, Ks, precision=prec)
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
        ms = e0.elapsed_time(e1) / args.steps
        mx, mean, viol = errors()
        info = plan.info()
        plan.close()
        r = {"gflops": flops / (ms * 1e-3) / 1e9, "ms_per_step": ms, "wall_s": wall,
             "max_rel_err": mx, "mean_rel_err": mean, "fp32_tol_violations_in_sample": viol,
             "work_items": info["work_items"],
             "hbm_gbs": nbytes / (ms * 1e-3) / 1e9}
        results[name] = r
        log(f"{name:7s}: {ms:.3f} ms/step  {r['gflops']:.0f} GFLOP/s  {r['hbm_gbs']:.0f} GB/s  "
            f"max_rel {mx:.3g} mean_rel {mean:.3g}  viol {viol}")
        if name == "bf16x3":
            headline = r

    # Dominant kernel: the 128 x 128-tile persistent bucket (GEMMs with
    # max(M, N) > 64).  A plan holding only those GEMMs is one kernel launch;
    # HIP events on its stream give that kernel's average duration.
    big = [g for g in range(args.count) if ls.query_batched_tile(ls.BF16X3, int(Ms[g]), int(Ns[g]), int(Ks[g]))
           == (128, 128)]
    bidx = np.array(big, dtype=np.int64)
    kplan = ls.BatchedPlan(0, 0, [Cs[g] for g in big], [As[g] for g in big], [Bs[g] for g in big],
                           Ms[bidx], Ns[bidx], Ks[bidx], precision=ls.BF16X3)
    kbytes = float(np.sum(4.0 * (Ms[bidx] * Ks[bidx] + Ks[bidx] * Ns[bidx] + Ms[bidx] * Ns[bidx])))
    kflops = float(np.sum(2.0 * Ms[bidx] * Ns[bidx] * Ks[bidx]))
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
    kms = k0.elapsed_time(k1) / args.steps
    kplan.close()
    kgbs = kbytes / (kms * 1e-3) / 1e9
    log(f"dominant kernel (128x128 bucket, {len(big)} GEMMs): {kms:.3f} ms/launch, {kgbs:.0f} GB/s, "
        f"{kflops / (kms * 1e-3) / 1e12:.1f} TFLOP/s")
    traffic = recorded_traffic()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(shapes, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    h = headline
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
        "dtype": "bf16x3 (fp32 in/out, bf16 MFMA, fp32 accumulate)",
        "data": "synthetic: U[-1,1) fp32 entries, M,N,K~U{1..128} splitmix64(0)",
        "config": {"workload": "batched SGEMM, 50k GEMMs, M,N,K~U{1..128}, NN, row-major",
                   "count": args.count, "gflop_per_step": round(flops / 1e9, 3),
                   "gb_per_step": round(nbytes / 1e9, 4)},
        "speedup_bf16x3_vs_fp32": (round(results["bf16x3"]["gflops"] / results["fp32"]["gflops"], 3)
                                   if "fp32" in results else None),
        "modes": {k: {kk: (float(f"{vv:.4g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(kgbs / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": "k_gemm_persistent<bf16x3, NN, 128x128> (the GEMMs with max(M,N) > 64)",
                     "kernel_ms": round(kms, 4), "kernel_algorithmic_bytes": kbytes,
                     "kernel_tflops": round(kflops / (kms * 1e-3) / 1e12, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def recorded_traffic():
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy) and recorded in
    profiles/; None when no such record exists."""
    path = os.path.join(ROOT, "profiles", "batched_sgemm_bf16x3_128_traffic.txt")
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


if __name__ == "__main__":
    main()
