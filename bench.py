#!/usr/bin/env python3
"""Benchmark of the Uni10 contraction hot path on one MI355X.

Workload (one "step"): the dense FP64 contraction-and-permute
  A(256,2,2,256) labels (0,1,2,3)  x  B(256,2,2,256) labels (3,2,4,5)
  -> contract over bonds {2,3} -> C(256,2,2,256) labels (0,1,4,5)
  -> permute C to positions (3,1,0,2) (labels 5,1,0,4)
with A, B ~ SplitMix64 uniform[-1,1) (seeds 1, 2), resident in HBM.  A step
runs exactly the device work of the drop-in contract() + permute(): the
permute of B to (shared | free), one 512x512x512 FP64 MFMA GEMM, the output
permute.

Reported value: GEMM TFLOP/s of the whole step (2*512^3 flop / step time).
Also reported: permute GB/s (16 B per element moved) at this shape and at an
HBM-resident (4096,2,2,4096) shape, per-kernel durations (HIP events on the
launch stream), the roofline of the dominant kernel, and the oracle's CPU
time for the same step on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
F64_MFMA_PEAK_TFLOPS = 78.6


def splitmix_uniform(seed, n):
    """SplitMix64 -> uniform[-1,1) (same stream as the oracle's generator)."""
    M = np.uint64
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = M(seed) + i * M(0x9E3779B97F4A7C15)
        z = (z ^ (z >> M(30))) * M(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> M(27))) * M(0x94D049BB133111EB)
        z = z ^ (z >> M(31))
    return 2.0 * ((z >> M(11)).astype(np.float64) * (2.0 ** -53)) - 1.0


def gpu_step_parts(tn, dims=(256, 2, 2, 256)):
    chi, d1, d2, chi2 = dims
    n = chi * d1 * d2 * chi2
    A = tn.DeviceArray.from_host(splitmix_uniform(1, n))
    B = tn.DeviceArray.from_host(splitmix_uniform(2, n))
    Bp = tn.DeviceArray((n,))
    C = tn.DeviceArray((n,))
    D = tn.DeviceArray((n,))
    M = chi * d1          # free bonds of A (labels 0,1)
    K = d2 * chi2         # shared bonds (labels 2,3)
    Nn = d1 * chi2        # free bonds of B (labels 4,5)
    desc = np.array([(M, Nn, K, 0, 0, 0, K, Nn, Nn)], dtype=tn.GEMM_DESC)
    plan = tn.GemmPlan(tn.TN_GEMM_F64, desc)

    def permute_b(stream=None):
        # B labels (3,2,4,5) -> (2,3,4,5): positions (1,0,2,3)
        tn.permute_dense(B, Bp, dims, (1, 0, 2, 3), stream)

    def gemm(stream=None):
        plan(A, Bp, C, stream)

    def permute_c(stream=None):
        tn.permute_dense(C, D, dims, (3, 1, 0, 2), stream)

    return dict(A=A, B=B, Bp=Bp, C=C, D=D, plan=plan, parts=[("permute_b", permute_b),
                                                                ("gemm_f64", gemm),
                                                                ("permute_c", permute_c)],
                flops=2.0 * M * Nn * K, n=n)


def time_parts(tn, parts, steps):
    """Per-part average duration with HIP events on the launch stream."""
    evs = {name: [] for name, _ in parts}
    for _ in range(steps):
        for name, fn in parts:
            e0, e1 = tn.Event(), tn.Event()
            e0.record()
            fn()
            e1.record()
            evs[name].append((e0, e1))
    tn.sync()
    return {name: float(np.mean([a.elapsed_ms(b) for a, b in pairs])) for name, pairs in evs.items()}


def permute_hbm(tn, steps):
    dims = (4096, 2, 2, 4096)
    n = int(np.prod(dims))
    src = tn.DeviceArray.from_host(splitmix_uniform(3, n))
    dst = tn.DeviceArray((n,))
    for _ in range(2):
        tn.permute_dense(src, dst, dims, (3, 1, 0, 2))
    tn.sync()
    e0, e1 = tn.Event(), tn.Event()
    e0.record()
    for _ in range(steps):
        tn.permute_dense(src, dst, dims, (3, 1, 0, 2))
    e1.record()
    ms = e0.elapsed_ms(e1) / steps
    src.free()
    dst.free()
    return 16.0 * n / (ms * 1e-3) / 1e9, ms


def cpu_baseline(budget_s=12.0, max_steps=400):
    """The oracle's CPU step: odometer permutes (uni10_cpu.c, one thread) and
    numpy dgemm (OpenBLAS) on the host cores."""
    from oracle import uni10_oracle as ora
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    cores = min(os.cpu_count() or 1, 16)
    dims = np.array([256, 2, 2, 256], np.int64)
    n = int(np.prod(dims))
    A = splitmix_uniform(1, n)
    B = splitmix_uniform(2, n)
    Bp = np.empty(n)
    D = np.empty(n)
    L = ora.lib()
    pb = np.array([1, 0, 2, 3], np.int32)
    pc = np.array([3, 1, 0, 2], np.int32)

    def step():
        L.ora_permute_dense(ora._ptr(B), ora._ptr(Bp), 4, ora._ptr(dims), ora._ptr(pb))
        C = np.ascontiguousarray(A.reshape(512, 512) @ Bp.reshape(512, 512)).reshape(-1)
        L.ora_permute_dense(ora._ptr(C), ora._ptr(D), 4, ora._ptr(dims), ora._ptr(pc))
        return C

    ctx = threadpool_limits(limits=cores) if threadpool_limits else None
    try:
        step()
        t0 = time.perf_counter()
        k = 0
        while k < max_steps and time.perf_counter() - t0 < budget_s:
            step()
            k += 1
        dt = (time.perf_counter() - t0) / k
    finally:
        if ctx is not None:
            ctx.unregister() if hasattr(ctx, "unregister") else None
    flops = 2.0 * 512 ** 3
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": "%d contraction-and-permute steps (256,2,2,256) FP64: C odometer permutes "
                      "(1 thread) + numpy/OpenBLAS dgemm (%d threads); %.2f ms/step" % (k, cores, dt * 1e3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent chains/requests would shard by instance; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from trim_amd import tn
    tn.init(0)
    info = tn.device_info()
    W = gpu_step_parts(tn)
    parts = W["parts"]

    def step():
        for _, fn in parts:
            fn()

    for _ in range(args.warmup):
        step()
    tn.sync()
    e0, e1 = tn.Event(), tn.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    tn.sync()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1) / args.steps
    value = W["flops"] / (ms * 1e-3) / 1e12

    per = time_parts(tn, parts, max(10, min(args.steps, 100)))
    gemm_ms = per["gemm_f64"]
    gemm_tflops = W["flops"] / (gemm_ms * 1e-3) / 1e12
    perm_ms = per["permute_c"]
    perm_gbps = 16.0 * W["n"] / (perm_ms * 1e-3) / 1e9
    hbm_gbps, hbm_ms = permute_hbm(tn, 10)
    dominant = max(per, key=per.get)
    if dominant == "gemm_f64":
        roof = {"bound": "mfma", "achieved": gemm_tflops, "peak": F64_MFMA_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": gemm_tflops / F64_MFMA_PEAK_TFLOPS, "traffic": None,
                "kernel": "gemm_mfma_16x16x4<double>", "avg_ms": gemm_ms}
    else:
        a = 16.0 * W["n"] / (per[dominant] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": a / HBM_PEAK_GBPS, "traffic": None, "kernel": dominant,
                "avg_ms": per[dominant]}
    out = {
        "metric": "sector-GEMM TFLOP/s + permute GB/s; TEBD steps/s, p99 latency vs plain FP64",
        "value": value, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dense FP64 contraction A(256,2,2,256) x B(256,2,2,256) over 2 bonds "
                               "+ permute (3,1,0,2)", "gemm": "512x512x512", "parallelism": "1 GPU"},
        "roofline": roof,
        "kernels_ms": per,
        "permute_GBps_config": perm_gbps,
        "permute_GBps_hbm_4096x2x2x4096": hbm_gbps,
        "permute_frac_hbm": hbm_gbps / HBM_PEAK_GBPS,
        "wall_s_timed": wall,
        "device": info,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
