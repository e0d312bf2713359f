#!/usr/bin/env python3
"""Benchmark of the Uni10 contraction hot path on one MI355X.

Headline workload (one "step"), FP64, through the drop-in C++ API:
    A(256,2,2,256) labels (0,1,2,3), B(256,2,2,256) labels (3,2,4,5)
    C = contract(A, B)                      -> (256,2,2,256) labels (0,1,4,5)
    C.permute([5,1,0,4], 2)                 -> positions (3,1,0,2)
A, B ~ SplitMix64 uniform[-1,1) (seeds 1, 2), resident in HBM before timing.
The step is exactly what a Uni10 caller runs; the device work is one
512x512x512 FP64 MFMA GEMM that reads A and B in their stored bond order and
one tiled permute.

value = 2*512^3 flop / step time (TFLOP/s of the whole step).

Also measured (same JSON line):
  * per-kernel durations with HIP events on the launch stream, and the
    roofline of the dominant kernel (GEMM: FP64 MFMA peak; permute: HBM);
  * permute GB/s (16 B per element) at this shape and at an HBM-resident
    (4096,2,2,4096) shape (512 MiB each way);
  * iTEBD steps/s, spin-1/2 Heisenberg, chi=32, d=2, tau=0.01 (a step is
    update(A,B) + update(B,A) through Networks, SVD and truncation);
  * the oracle's CPU time for the same contraction step and iTEBD step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
NETDIR = os.path.join(ROOT, "trim_amd", "nets")     # the network files shipped with the package

HBM_PEAK_GBPS = 8000.0
F64_MFMA_PEAK_TFLOPS = 78.6
DIMS = (256, 2, 2, 256)
# FP64 GEMM tile shapes by autotuner id: BM x BN block tile, KC k per stage
F64_TILES = {0: "32x32,kc32", 1: "64x64,kc64", 2: "128x64,kc32", 3: "64x128,kc32", 4: "64x32,kc64",
             5: "32x32,kc32,512 threads", 6: "32x32,kc64,512 threads", 7: "32x32,kc128,512 threads",
             8: "32x32,kc32,2 stages in flight"}


def splitmix_uniform(seed, n):
    """SplitMix64 -> uniform[-1,1): x_i = mix(seed + i*0x9E3779B97F4A7C15), i >= 1."""
    M = np.uint64
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = M(seed) + i * M(0x9E3779B97F4A7C15)
        z = (z ^ (z >> M(30))) * M(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> M(27))) * M(0x94D049BB133111EB)
        z = z ^ (z >> M(31))
    return 2.0 * ((z >> M(11)).astype(np.float64) * (2.0 ** -53)) - 1.0


def make_tensors(u10):
    n = int(np.prod(DIMS))
    bonds = [u10.Bond(u10.BD_IN, DIMS[0]), u10.Bond(u10.BD_IN, DIMS[1]),
             u10.Bond(u10.BD_OUT, DIMS[2]), u10.Bond(u10.BD_OUT, DIMS[3])]
    A = u10.UniTensor(bonds, "A")
    B = u10.UniTensor(bonds, "B")
    A.setElem(splitmix_uniform(1, n))
    B.setElem(splitmix_uniform(2, n))
    A.setLabel([0, 1, 2, 3])
    B.setLabel([3, 2, 4, 5])
    return A, B


def api_step(u10, A, B):
    C = u10.contract(A, B)
    C.permute([5, 1, 0, 4], 2)
    return C


def kernel_parts(tn):
    """The device work of one step, issued directly on the C ABI with the same
    plans the API builds (gather-addressed GEMM, tiled permute)."""
    chi, d1, d2, chi2 = DIMS
    n = chi * d1 * d2 * chi2
    A = tn.DeviceArray.from_host(splitmix_uniform(1, n))
    B = tn.DeviceArray.from_host(splitmix_uniform(2, n))
    C = tn.DeviceArray((n,))
    D = tn.DeviceArray((n,))
    # A labels (0,1,2,3), strides (1024,512,256,1): rows (0,1) and K (2,3)
    # are each one contiguous level.  B labels (3,2,4,5), same strides:
    # K = (label 2, label 3) -> levels (2, 512), (256, 1024); N = (4,5) contiguous.
    M, K, N = chi * d1, d2 * chi2, d1 * chi2
    desc = np.zeros(1, dtype=tn.GEMM_DESC_EX)
    desc["M"], desc["N"], desc["K"] = M, N, K
    desc["lda"], desc["ldb"], desc["ldc"] = K, N, N
    desc["a_row"] = tn.index_map([(M, K)])
    desc["a_col"] = tn.index_map([(K, 1)])
    desc["b_row"] = tn.index_map([(d2, d1 * chi2), (chi, d2 * d1 * chi2)])
    desc["b_col"] = tn.index_map([(N, 1)])
    plan = tn.GemmPlan(tn.TN_GEMM_F64, desc)
    plan.autotune(A, B, C, reps=3)

    def gemm(stream=None):
        plan(A, B, C, stream)

    def permute(stream=None):
        tn.permute_dense(C, D, DIMS, (3, 1, 0, 2), stream)

    info = plan.info()
    keep = (A, B, C, D, desc, plan)
    return [("gemm_f64", gemm), ("permute", permute)], 2.0 * M * N * K, n, keep, \
        {"gemm_tiles": info["tiles"], "gemm_tile_cfg": info["tile_id"]}


def time_parts(tn, parts, steps, pairs=None):
    """Average launch duration of each part, with HIP events on the launch
    stream.  A few large fill kernels are queued first so the device is
    still busy while the host enqueues: the events then see device time, not
    host enqueue gaps.  The figure returned is one event pair around `steps`
    back-to-back launches, divided by `steps` (kernel time plus the device's
    dispatch gap between dependent launches).  With `pairs` (a dict), the
    mean over an event pair around every single launch is stored there too
    (it adds the events' own cost to each launch)."""
    spacer = tn.DeviceArray((256 * 1024 * 1024,))   # 2 GiB
    out = {}
    for name, fn in parts:
        for _ in range(3):
            fn()
        tn.sync()
        for _ in range(max(4, steps // 25)):
            tn.call("tn_zero", tn._p(spacer), spacer.nbytes, None)
        e0, e1 = tn.Event(), tn.Event()
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        tn.sync()
        out[name] = e0.elapsed_ms(e1) / steps
        if pairs is not None:
            evs = [(tn.Event(), tn.Event()) for _ in range(steps)]
            for _ in range(max(4, steps // 25)):
                tn.call("tn_zero", tn._p(spacer), spacer.nbytes, None)
            for a, b in evs:
                a.record()
                fn()
                b.record()
            tn.sync()
            pairs[name] = float(np.mean([a.elapsed_ms(b) for a, b in evs]))
    spacer.free()
    return out


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


def copy_rate(tn, steps=10):
    """Plain device-to-device copy of the same 1 GiB moved by the HBM-shape
    permute (512 MiB read + 512 MiB written): what the library's copy path
    reaches on this HBM, the practical ceiling the permute is set against."""
    n = 4096 * 2 * 2 * 4096
    a = tn.DeviceArray((n,), zero=True)
    b = tn.DeviceArray((n,))
    for _ in range(2):
        tn.call("tn_memcpy_d2d", tn._p(b), tn._p(a), a.nbytes, None)
    tn.sync()
    e0, e1 = tn.Event(), tn.Event()
    e0.record()
    for _ in range(steps):
        tn.call("tn_memcpy_d2d", tn._p(b), tn._p(a), a.nbytes, None)
    e1.record()
    ms = e0.elapsed_ms(e1) / steps
    a.free()
    b.free()
    return 16.0 * n / (ms * 1e-3) / 1e9


def copy_calibration(tn, mib=256):
    """One plain device-to-device copy of a known size: the reference the
    FETCH_SIZE / WRITE_SIZE counters are scaled by (bytes the counter reports
    per byte actually moved)."""
    n = mib * 1024 * 1024 // 8
    a = tn.DeviceArray((n,), zero=True)
    b = tn.DeviceArray((n,))
    tn.call("tn_memcpy_d2d", tn._p(b), tn._p(a), a.nbytes, None)
    tn.sync()
    a.free()
    b.free()
    return mib * 1024 * 1024


def recorded_traffic(kernel):
    """HBM bytes per launch of the headline GEMM from the committed counter
    summary (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, scaled by
    the calibration copy), or None when there is none for this kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_gemm_f64_512x512x512.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("kernels", {}).get(kernel)


def c3_sector_operands(tn, u10, A1, A2):
    """The sector GEMMs of A1 . A2 (one per middle-bond charge): A1 permuted
    to (left, phys | middle), A2 to (middle | phys, right); descriptors for
    one grouped launch."""
    A1p = A1.copy().permute([-1, 10, 0], 2)
    A2p = A2.copy().permute([0, 11, -2], 1)
    b1 = {q.triple(): (r, c, o) for q, r, c, o in A1p.blocks()}
    b2 = {q.triple(): (r, c, o) for q, r, c, o in A2p.blocks()}
    descs, oc = [], 0
    for q, (r1, c1, o1) in b1.items():
        if q in b2:
            r2, c2, o2 = b2[q]
            descs.append((r1, c2, c1, o1, o2, oc, max(c1, 1), max(c2, 1), max(c2, 1)))
            oc += r1 * c2
    C = tn.DeviceArray((oc,))
    return descs, A1p.device_ptr(), A2p.device_ptr(), C, (A1p, A2p)


def c3_sector_only(tn, u10, steps, mode=0):
    """Only the grouped A1 . A2 sector GEMM of the U(1) chi=1024 workload:
    autotune, then `steps` timed launches (for kernel-trace and counter
    runs)."""
    from trim_amd import workloads
    A1, A2, G, dims = workloads.u1_two_site(u10, 1024, 8, 11, 0.01)
    descs, pa, pb, C, keep = c3_sector_operands(tn, u10, A1, A2)
    plan = tn.GemmPlan(mode, np.array(descs, dtype=tn.GEMM_DESC))
    plan.autotune(pa, pb, C, reps=3)
    info = plan.info()
    pairs = {}
    t = time_parts(tn, [("sector", lambda: plan(pa, pb, C))], steps, pairs)["sector"]
    per_cfg = {}
    for cfg in range(9 if mode in (0, 1) else (4 if mode == 4 else 5)):
        plan.set_tile(cfg)
        per_cfg[cfg] = time_parts(tn, [("c", lambda: plan(pa, pb, C))], steps, {})["c"]
    plan.set_tile(info["tile_id"])
    return {"sectors": len(descs), "mode": mode, "tiles": info["tiles"], "tile_cfg": info["tile_id"],
            "kernel_ms_per_tile_cfg": per_cfg,
            "flop": info["flops"], "kernel_ms": t, "kernel_ms_event_pairs": pairs["sector"],
            "tflops": info["flops"] / (t * 1e-3) / 1e12,
            "shapes_MNK": [(d[0], d[1], d[2]) for d in descs]}


def c3_u1(tn, u10, steps=10):
    """U(1) two-site contraction, chi=1024 over Sz in {-8..8} (binomial sector
    dims), d=2: Theta = (A1 . A2) . G through the drop-in API in every GEMM
    mode, plus the sector GEMM of A1 . A2 alone as one grouped launch."""
    from trim_amd import workloads
    A1, A2, G, dims = workloads.u1_two_site(u10, 1024, 8, 11, 0.01)
    res = {"sector_dims": dims}
    ref = None
    modes = (("f64", 0), ("f32", 1), ("bf16x3", 2), ("bf16x6", 3), ("bf16_exact", 4))
    for name, mode in modes:
        u10.set_gemm_mode(mode)
        R = u10.contract(u10.contract(A1, A2), G)      # plans + autotune
        tn.call("tn_device_sync")
        e0, e1 = tn.Event(), tn.Event()
        e0.record()
        for _ in range(steps):
            R = u10.contract(u10.contract(A1, A2), G)
        e1.record()
        ms = e0.elapsed_ms(e1) / steps
        out = R.getElem()
        if ref is None:
            ref = out
        err = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
        res[name] = {"ms_per_contraction": ms, "rel_frobenius_err_vs_f64": err}
    u10.set_gemm_mode(0)
    # the sector GEMM of A1 . A2 as one grouped launch
    descs, pa, pb, C, keep = c3_sector_operands(tn, u10, A1, A2)
    sect = {"sectors": len(descs)}
    for name, mode in modes:
        plan = tn.GemmPlan(mode, np.array(descs, dtype=tn.GEMM_DESC))
        plan.autotune(pa, pb, C, reps=3)
        info = plan.info()
        t = time_parts(tn, [(name, lambda: plan(pa, pb, C))], 50)[name]
        sect[name] = {"kernel_ms": t, "tflops": info["flops"] / (t * 1e-3) / 1e12, "tiles": info["tiles"],
                      "tile_cfg": info["tile_id"], "flop": info["flops"]}
    res["sector_gemm_A1A2"] = sect
    return res


def tebd_rate(u10, steps, chi=32, graph=False):
    t = u10.ITebd(chi, 2, 1.0, 0.01, 7, NETDIR)
    t.set_graph(graph)
    t.step(5)
    t0 = time.perf_counter()
    t.step(steps)
    dt = time.perf_counter() - t0
    return steps / dt, t.energy(), t.trunc_err()


def c4_batch(u10, chis, steps, label, u1_half=False):
    """Batched iTEBD: 512 spin-1/2 XXZ chains (delta ~ U[0,2]) with bond
    dimensions cycling through chis, every update stage one launch for all
    chains, steps replayed as a captured graph.  u1_half: chains 0-255 U(1),
    256-511 dense.  One step = update(A,B) + update(B,A) of every chain."""
    n = 512
    chi = [int(chis[i % len(chis)]) for i in range(n)]
    sym = [1 if (u1_half and i < n // 2) else 0 for i in range(n)]
    delta = np.random.default_rng(4).uniform(0.0, 2.0, n)
    B = u10.BatchTebd(chi, delta, [1000 + 4 * i for i in range(n)], 0.01, NETDIR, sym=sym)
    B.set_graph(True)
    B.step(1)   # first SVDs from scratch
    B.step(1)   # capture
    t0 = time.perf_counter()
    B.step(steps)
    dt = time.perf_counter() - t0
    info = B.info()
    return {"chains": n, "chi": label, "symmetry": "chains 0-255 U(1), 256-511 dense" if u1_half else "dense",
            "steps": steps, "ms_per_step": dt / steps * 1e3,
            "steps_per_s": steps / dt, "chain_updates_per_s": 2.0 * n * steps / dt,
            "gemm_tflops": info["gemm_flops_per_step"] * steps / dt / 1e12,
            "svd_sweeps_last_update": info["last_sweeps"]}


def large_chi_steps(u10):
    """One chain per (chi, d, symmetry) at the top of the serving range: time
    of the first (cold-SVD) step and of a warm step."""
    out = {}
    for chi in (256, 512, 1024):
        for d in (2, 3):
            for sym in (0, 1):
                B = u10.BatchTebd([chi], [1.0], [7], 0.01, NETDIR, d=[d], sym=[sym])
                t0 = time.perf_counter()
                B.step(1)
                t1 = time.perf_counter()
                B.step(1)
                t2 = time.perf_counter()
                out["chi%d_d%d_%s" % (chi, d, "u1" if sym else "dense")] = {
                    "first_step_ms": (t1 - t0) * 1e3, "warm_step_ms": (t2 - t1) * 1e3,
                    "svd_sweeps": B.info()["last_sweeps"]}
                progress("chi=%d d=%d %s: warm step %.1f ms" % (chi, d, "U(1)" if sym else "dense", (t2 - t1) * 1e3))
                del B
    return out


def mixed_precision_energy(u10, checkpoints=(20, 2000)):
    """Energy per site of four chains after each checkpoint's number of
    batched steps in each GEMM mode, against the FP64 run (SVD, truncation
    and energies FP64 in every mode): early in the evolution, where the
    energy still moves at first order, and near the converged ground state,
    where it is stationary."""
    chi, delta, seed = [8, 16, 24, 32], [0.5, 1.0, 1.5, 1.0], [21, 22, 23, 24]
    e = {c: {} for c in checkpoints}
    last_move = {}
    for mode, name in ((0, "f64"), (1, "f32"), (2, "bf16x3"), (3, "bf16x6"), (4, "bf16_exact")):
        B = u10.BatchTebd(chi, delta, seed, 0.01, NETDIR, mode)
        B.set_graph(True)
        done = 0
        for c in checkpoints:
            if c - done > 100 and c == checkpoints[-1]:
                B.step(c - 100 - done)
                e0 = np.array([B.energy(i) for i in range(len(chi))])
                done = c - 100
                B.step(c - done)
                e[c][name] = np.array([B.energy(i) for i in range(len(chi))])
                last_move[name] = float(np.max(np.abs(e[c][name] - e0)))
            else:
                B.step(c - done)
                e[c][name] = np.array([B.energy(i) for i in range(len(chi))])
            done = c
        del B
    out = {}
    for c in checkpoints:
        out["after_%d_steps" % c] = {k: float(np.max(np.abs(v - e[c]["f64"]))) for k, v in e[c].items() if k != "f64"}
    out["f64_energy_change_over_last_100_steps"] = last_move.get("f64")
    return out


def serve_budget():
    """The serving memory budget: 80 % of the device's HBM (288 GB on one
    MI355X) beyond what the process holds already."""
    from trim_amd import tn
    return 0.8 * tn.device_info()["hbm_bytes"]


def c5_serve(u10, n_req, rate, chis=(32, 64, 128), mixed=True):
    """Poisson stream of TEBD-step and contraction requests (bond dimension
    log-uniform over chis, 1..50 steps; mixed: d in {2, 3} and 60 % U(1),
    else dense spin 1/2) served by continuous batching over resident slots,
    FP64, FP32-MFMA and exact split-BF16 GEMMs, every request served to the
    end."""
    from trim_amd import serve
    if mixed:
        classes = [(c, d, s) for c in chis for d in (2, 3) for s in (0, 1)]
        gen = dict(d_choices=(2, 3), u1_fraction=0.6)
    else:
        classes = [(c, 2, 0) for c in chis]
        gen = {}
    out, res = {}, {}
    probe = None
    for mode, name in ((0, "f64"), (1, "f32"), (4, "bf16_exact")):
        srv = serve.Server(classes, serve_budget(), NETDIR, gemm_mode=mode, max_slots=256,
                           slots_per_engine=64 if mixed else 256, step_ms=probe)
        probe = srv.step_ms
        reqs = serve.poisson_requests(n_req, rate, seed=5, chi_classes=list(chis), **gen)
        st = srv.run(reqs, realtime=True)
        res[name] = {r.rid: r.result for r in st["done"]}
        out[name] = {k: st[k] for k in ("requests", "failed", "wall_s", "requests_per_s", "p50_ms", "p99_ms",
                                        "tebd_steps_per_s", "chain_steps_per_s", "engines", "lanes", "host_s")}
        out[name]["slots"] = {"chi%d_d%d_%s" % (c, d, "u1" if s else "dense"): srv.slots((c, d, s))
                              for c, d, s in classes}
        srv.close()
        del srv
        progress("serving %s: %.0f req/s, p99 %.1f ms" % (name, out[name]["requests_per_s"], out[name]["p99_ms"]))
    diff = max(float(np.max(np.abs(res["f32"][k] - res["f64"][k]))) for k in res["f64"])
    out["max_result_diff_f32_vs_f64"] = diff
    out["max_result_diff_bf16_exact_vs_f64"] = max(float(np.max(np.abs(res["bf16_exact"][k] - res["f64"][k])))
                                                   for k in res["f64"])
    out["workload"] = ("Poisson %.0f req/s, %d requests, chi log-uniform over %s, %s, 1..50 steps, half TEBD "
                       "half contraction, served to the end" %
                       (rate, n_req, list(chis), "d in {2,3}, 60% U(1) (charges -12..12)" if mixed
                        else "dense spin 1/2"))
    return out


SERVE_RANGE = (32, 64, 128, 256, 512, 1024)


def c5_serve_window(u10, rate, window_s, modes=((0, "f64"), (4, "bf16_exact"))):
    """The serving workload over the whole bond-dimension range 32..1024
    (classes 32, 64, ..., 1024; d in {2, 3}; 60 % U(1); 1..50 steps; half
    TEBD, half contraction), Poisson at `rate`, for a fixed serving window:
    requests that arrive in the window are admitted as slots and the budget
    allow; at its end the work on the device finishes and the rest counts
    as unfinished.  Engines start at 8 slots and grow on demand."""
    from trim_amd import serve
    classes = [(c, d, s) for c in SERVE_RANGE for d in (2, 3) for s in (0, 1)]
    n = int(rate * window_s * 1.05) + 16
    out = {}
    probe = None
    for mode, name in modes:
        t = time.perf_counter()
        srv = serve.Server(classes, serve_budget(), NETDIR, gemm_mode=mode, max_slots=256, slots_per_engine=64,
                           step_ms=probe)
        probe = srv.step_ms
        t_setup = time.perf_counter() - t
        reqs = serve.poisson_requests(n, rate, seed=21, chi_classes=list(SERVE_RANGE), d_choices=(2, 3),
                                      u1_fraction=0.6)
        st = srv.run(reqs, realtime=True, deadline_s=window_s)
        # the TEBD work offered in the window, in seconds of one-chain steps
        # (each class's measured step time): how far beyond one device the
        # offered load is
        offered_work = sum(r.steps * srv.step_ms[(r.chi, r.d, r.sym)] * 1e-3 for r in reqs
                           if r.kind == "tebd" and r.t_arrival < window_s)
        done = [r for r in st["done"] if r.error is None]
        by_chi = {}
        for c in SERVE_RANGE:
            for kind in ("tebd", "contract"):
                lat = [r.t_done - r.t_arrival for r in done if r.chi == c and r.kind == kind]
                by_chi["chi%d_%s" % (c, kind)] = {
                    "completed": len(lat), "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat else None}
        out[name] = {"offered": st["offered"], "completed": st["requests"] - st["failed"], "failed": st["failed"],
                     "unfinished": st["unfinished"], "window_s": window_s, "wall_s": st["wall_s"],
                     "completed_per_s": (st["requests"] - st["failed"]) / window_s,
                     "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                     "chain_steps_per_s": st["chain_steps_per_s"], "engines": st["engines"],
                     "setup_s": t_setup, "host_s": st["host_s"], "lanes": st["lanes"],
                     "offered_tebd_one_chain_step_s_per_s": offered_work / window_s,
                     "step_ms_one_chain": {"chi%d_d%d_%s" % (c[0], c[1], "u1" if c[2] else "dense"): v
                                           for c, v in srv.step_ms.items()},
                     "by_chi": by_chi}
        srv.close()
        del srv
        progress("serving window %s: %d of %d offered completed, p99 %s ms" %
                 (name, out[name]["completed"], st["offered"], out[name]["p99_ms"]))
    out["workload"] = ("Poisson %.0f req/s offered for a %.0f s window, chi log-uniform over %s, d in {2,3}, "
                       "60%% U(1), 1..50 steps, half TEBD half contraction" % (rate, window_s, list(SERVE_RANGE)))
    return out


LANCZOS_N = 8192
LANCZOS_STEPS = 60


def lanczos_matrix(n):
    """Symmetric X + X^T, X ~ U[-1, 1) (seed 21): the Lanczos leg's input."""
    X = np.random.default_rng(21).uniform(-1.0, 1.0, (n, n))
    X += X.T.copy()
    return X


def lanczos_leg(tn):
    """lanczosEigh on the device: a fixed number of Lanczos steps (err_tol 0,
    so the run never stops early) on an 8192 x 8192 symmetric matrix (512
    MiB in HBM).  Every step reads the matrix once (matrix-vector product),
    so n^2 * 8 bytes per step over the step time is the HBM rate the step
    reaches, including its vector work and the two doubles copied back."""
    import ctypes
    n = LANCZOS_N
    A = tn.DeviceArray.from_host(lanczos_matrix(n))
    start = np.random.default_rng(22).uniform(-1.0, 1.0, n)
    res = {}
    for label, steps in (("warm", 5), ("timed", LANCZOS_STEPS)):
        psi = tn.DeviceArray.from_host(start)
        it = ctypes.c_int64(steps)
        e0 = ctypes.c_double()
        conv = ctypes.c_int()
        tn.sync()
        t0 = time.perf_counter()
        tn.call("tn_lanczos_f64", tn._p(A), n, tn._p(psi), ctypes.byref(it), 0.0, ctypes.byref(e0),
                ctypes.byref(conv), None)
        dt = time.perf_counter() - t0
        res[label] = (dt, it.value, e0.value)
    dt, steps, e0 = res["timed"]
    per = dt / steps
    gbps = n * n * 8 / per / 1e9
    del A
    return {"n": n, "steps": steps, "ms_per_step": per * 1e3, "lowest_ritz_value": e0,
            "matrix_GBps_per_step": gbps, "frac_hbm": gbps / HBM_PEAK_GBPS,
            "note": "matrix bytes per step / step time (whole step: matvec, vector work, host bisection)"}


def cpu_threads():
    """Threads the host leg uses for 'all cores': the CPUs this process may
    run on, capped by OMP_NUM_THREADS when the environment sets it (the
    machine's share for this job)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(avail, omp) if omp > 0 else avail), avail


def cpu_baseline(budget_s=8.0, max_steps=400):
    """The oracle's contraction step on the host: C odometer permutes (one
    thread) and numpy/OpenBLAS dgemm, timed with one BLAS thread and with
    every thread available to the process."""
    from oracle import uni10_oracle as ora
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    cores, avail = cpu_threads()
    dims = np.array(DIMS, np.int64)
    n = int(np.prod(dims))
    A = splitmix_uniform(1, n)
    B = splitmix_uniform(2, n)
    Bp, Bb, D = np.empty(n), np.empty(n), np.empty(n)
    L = ora.lib()
    pb = np.array([1, 0, 2, 3], np.int32)
    pc = np.array([3, 1, 0, 2], np.int32)

    def step():
        # contract(): permute B to (shared | free), dgemm, permute B back; then permute C
        L.ora_permute_dense(ora._ptr(B), ora._ptr(Bp), 4, ora._ptr(dims), ora._ptr(pb))
        C = np.ascontiguousarray(A.reshape(512, 512) @ Bp.reshape(512, 512)).reshape(-1)
        L.ora_permute_dense(ora._ptr(Bp), ora._ptr(Bb), 4, ora._ptr(dims), ora._ptr(pb))
        L.ora_permute_dense(ora._ptr(C), ora._ptr(D), 4, ora._ptr(dims), ora._ptr(pc))

    def timed(threads):
        lim = threadpool_limits(limits=threads) if threadpool_limits else None
        try:
            step()
            t0 = time.perf_counter()
            k = 0
            while k < max_steps and time.perf_counter() - t0 < budget_s:
                step()
                k += 1
            return (time.perf_counter() - t0) / k, k
        finally:
            if lim is not None and hasattr(lim, "unregister"):
                lim.unregister()
    dt1, k1 = timed(1)
    dt, k = timed(cores)
    # iTEBD on the oracle (chi=32), a short sample
    t = ora.ITEBD(32, 2, ora.heisenberg_two_site(1.0), 0.01, 7, NETDIR)
    t.step()
    t1 = time.perf_counter()
    ns = 0
    while ns < 200 and time.perf_counter() - t1 < 5.0:
        t.step()
        ns += 1
    tebd = ns / (time.perf_counter() - t1)
    # Lanczos on the oracle (numpy dgemv, BLAS threads = cores), a few steps
    lim = threadpool_limits(limits=cores) if threadpool_limits else None
    S = lanczos_matrix(LANCZOS_N)
    start = np.random.default_rng(22).uniform(-1.0, 1.0, LANCZOS_N)
    ora.lanczos_ev(S, start, max_iter=3, err_tol=0.0)
    t2 = time.perf_counter()
    _, _, lz_steps, _ = ora.lanczos_ev(S, start, max_iter=20, err_tol=0.0)
    lz = (time.perf_counter() - t2) / lz_steps
    del S
    if lim is not None and hasattr(lim, "unregister"):
        lim.unregister()
    return {"value": 2.0 * 512 ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": "%d contraction-and-permute steps (256,2,2,256) FP64 in %.1f s: oracle C odometer "
                      "permutes (1 thread) + numpy/OpenBLAS dgemm (%d threads), %.3f ms/step (1 thread: %d steps, "
                      "%.3f ms/step); oracle iTEBD chi=32: %d steps, %.1f steps/s"
                      % (k, k * dt, cores, dt * 1e3, k1, dt1 * 1e3, ns, tebd),
            "value_1_thread": 2.0 * 512 ** 3 / dt1 / 1e12, "ms_per_step": dt * 1e3, "ms_per_step_1_thread": dt1 * 1e3,
            "nproc": os.cpu_count(), "cpus_available": avail,
            "tebd_steps_per_s": tebd, "lanczos_ms_per_step_n8192": lz * 1e3}


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--tebd-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-chain and serving legs")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-window", type=float, default=15.0,
                    help="seconds of the serving window over chi 32..1024")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving legs, one JSON line")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline step, its kernels and the copy calibration (for counter runs)")
    ap.add_argument("--batch-only", type=int, default=0, metavar="STEPS",
                    help="only the 512-chain chi 16..256 half-U(1) batch for STEPS steps, one JSON line "
                         "(for kernel-trace runs of the batched update)")
    ap.add_argument("--c3-only", type=int, default=0, metavar="STEPS",
                    help="only the grouped U(1) chi=1024 A1.A2 sector GEMM, STEPS timed launches, one JSON line "
                         "(for kernel-trace and counter runs)")
    ap.add_argument("--c3-mode", type=int, default=0, help="GEMM mode of --c3-only (0 FP64)")
    ap.add_argument("--tebd-only", type=int, default=0, metavar="STEPS",
                    help="only the single-chain chi=32 iTEBD through the drop-in API for STEPS steps, one JSON "
                         "line (for kernel-trace runs)")
    ap.add_argument("--lanczos-only", action="store_true",
                    help="only the Lanczos leg (8192 x 8192), one JSON line (for kernel-trace runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path in this build (independent "
              "chains and requests would shard by instance across GPUs). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    from trim_amd import tn, uni10
    tn.init(0)
    uni10.load()
    info = tn.device_info()
    if args.tebd_only:
        sps, e, te = tebd_rate(uni10, args.tebd_only)
        print(json.dumps({"metric": "iTEBD steps/s (chi=32, d=2, drop-in API)", "value": sps, "unit": "steps/s",
                          "energy": e, "trunc_err": te}), flush=True)
        return
    if args.lanczos_only:
        print(json.dumps(lanczos_leg(tn)), flush=True)
        return
    if args.c3_only:
        print(json.dumps(c3_sector_only(tn, uni10, args.c3_only, args.c3_mode)), flush=True)
        return
    if args.serve_only:
        out = c5_serve(uni10, args.serve_requests, 2000.0)
        progress("serving done")
        out["chi_32_to_1024_window"] = c5_serve_window(uni10, 2000.0, args.serve_window)
        out["large_chi_single_chain_steps"] = large_chi_steps(uni10)
        print(json.dumps({"metric": "serving (Poisson 2000 req/s)", "detail": out}), flush=True)
        return
    if args.batch_only:
        out = c4_batch(uni10, list(range(16, 257, 16)), args.batch_only, "16..256 step 16, 32 chains each",
                       u1_half=True)
        print(json.dumps({"metric": "batched iTEBD steps/s (512 chains, chi 16..256, half U(1))",
                          "value": out["steps_per_s"], "unit": "steps/s", "detail": out}))
        return

    # headline: the drop-in API step
    A, B = make_tensors(uni10)
    for _ in range(args.warmup):
        api_step(uni10, A, B)
    tn.call("tn_device_sync")
    e0, e1 = tn.Event(), tn.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        api_step(uni10, A, B)
    e1.record()
    tn.call("tn_device_sync")
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1) / args.steps
    flops = 2.0 * 512 ** 3
    value = flops / (ms * 1e-3) / 1e12

    progress("headline step done: %.4f ms/step" % ms)
    parts, kflops, n, keep, ginfo = kernel_parts(tn)
    pairs = {}
    per = time_parts(tn, parts, max(20, min(args.steps, 200)), pairs)
    gemm_tf = kflops / (per["gemm_f64"] * 1e-3) / 1e12
    perm_gbps = 16.0 * n / (per["permute"] * 1e-3) / 1e9
    hbm_gbps, hbm_ms = permute_hbm(tn, 10)
    copy_gbps = copy_rate(tn)
    if per["gemm_f64"] >= per["permute"]:
        roof = {"bound": "mfma", "achieved": gemm_tf, "peak": F64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": gemm_tf / F64_MFMA_PEAK_TFLOPS, "traffic": None,
                "kernel": "gemm_mfma_k32<double,%s>" % F64_TILES.get(ginfo["gemm_tile_cfg"], "?"),
                "avg_ms": per["gemm_f64"],
                "flop_per_launch": kflops}
    else:
        roof = {"bound": "hbm", "achieved": perm_gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": perm_gbps / HBM_PEAK_GBPS, "traffic": None, "kernel": "tiled_permute_kernel",
                "avg_ms": per["permute"], "bytes_per_launch": 16 * n}
    copy_bytes = copy_calibration(tn)
    roof["traffic"] = recorded_traffic(roof["kernel"])
    progress("kernel parts and HBM permute done")
    if args.headline_only:
        print(json.dumps({"metric": "headline step (counter run)", "ms_per_step": ms, "roofline": roof,
                          "kernels_ms": per, "kernels_ms_event_pairs": pairs, "copy_calibration_bytes": copy_bytes,
                          "permute_GBps_hbm_4096x2x2x4096": hbm_gbps, "copy_GBps_same_bytes": copy_gbps}),
              flush=True)
        return
    tebd_sps, tebd_e, tebd_te = tebd_rate(uni10, args.tebd_steps)
    progress("iTEBD chi=32: %.1f steps/s" % tebd_sps)
    tebd_g_sps, tebd_g_e, _ = tebd_rate(uni10, args.tebd_steps, graph=True)
    progress("iTEBD chi=32, networks replayed as graphs: %.1f steps/s" % tebd_g_sps)
    c3 = c3_u1(tn, uni10)
    progress("U(1) chi=1024 contraction modes done")
    lz = lanczos_leg(tn)
    progress("Lanczos n=%d: %.3f ms/step" % (lz["n"], lz["ms_per_step"]))
    c4 = {}
    if not args.no_batch:
        c4["chi_16_32"] = c4_batch(uni10, [16, 32], 10, "16 and 32, 256 chains each")
        progress("batched iTEBD chi 16/32: %.1f steps/s" % c4["chi_16_32"]["steps_per_s"])
        c4["chi_16_to_256_half_u1"] = c4_batch(uni10, list(range(16, 257, 16)), 3,
                                               "16..256 step 16, 32 chains each", u1_half=True)
        progress("batched iTEBD chi 16..256 half U(1): %.3f steps/s" % c4["chi_16_to_256_half_u1"]["steps_per_s"])
        c4["chi_16_to_256_dense"] = c4_batch(uni10, list(range(16, 257, 16)), 3, "16..256 step 16, 32 chains each")
        progress("batched iTEBD chi 16..256 dense: %.3f steps/s" % c4["chi_16_to_256_dense"]["steps_per_s"])
        c4["energy_err_vs_f64"] = mixed_precision_energy(uni10)
        progress("mixed-precision energies done")
        c5 = c5_serve(uni10, args.serve_requests, 2000.0)
        progress("serving done")
        c5["chi_32_to_1024_window"] = c5_serve_window(uni10, 2000.0, args.serve_window)
        progress("serving window over chi 32..1024 done")
        c5["spin_half_dense_chi_16_32"] = c5_serve(uni10, args.serve_requests, 2000.0, chis=(16, 32), mixed=False)
        progress("spin-1/2 dense serving done")
        c5["large_chi_single_chain_steps"] = large_chi_steps(uni10)
        c5["speedup_f32_serving_vs_f64_dropin"] = c5["f32"]["tebd_steps_per_s"] / tebd_sps
        c5["speedup_bf16_exact_serving_vs_f64_dropin"] = c5["bf16_exact"]["tebd_steps_per_s"] / tebd_sps
        c5["speedup_f64_serving_vs_f64_dropin"] = c5["f64"]["tebd_steps_per_s"] / tebd_sps
        c5["spin_half_dense_chi_16_32"]["speedup_f32_serving_vs_f64_dropin"] = \
            c5["spin_half_dense_chi_16_32"]["f32"]["tebd_steps_per_s"] / tebd_sps
    out = {
        "metric": "sector-GEMM TFLOP/s + permute GB/s; TEBD steps/s, p99 latency vs plain FP64",
        "value": value, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dense FP64 contract(A(256,2,2,256), B(256,2,2,256)) over 2 bonds + "
                               "permute (3,1,0,2), drop-in API", "gemm": "512x512x512",
                   "parallelism": "1 GPU"},
        "roofline": roof,
        "kernels_ms": per,
        "kernels_ms_event_pairs": pairs,
        "gemm_plan": ginfo,
        "permute_GBps_config": perm_gbps,
        "permute_GBps_hbm_4096x2x2x4096": hbm_gbps,
        "permute_frac_hbm": hbm_gbps / HBM_PEAK_GBPS,
        "copy_GBps_same_bytes": copy_gbps,
        "permute_frac_of_copy": hbm_gbps / copy_gbps,
        "tebd_chi32": {"steps_per_s": tebd_sps, "steps": args.tebd_steps, "energy_per_site": tebd_e,
                       "trunc_err_last": tebd_te, "steps_per_s_network_graphs": tebd_g_sps,
                       "energy_per_site_network_graphs": tebd_g_e},
        "c3_u1_chi1024": c3,
        "lanczos": lz,
        "c4_batched_itebd": c4,
        "c5_serving": (c5 if not args.no_batch else None),
        "wall_s_timed": wall,
        "device": info,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        progress("cpu baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
