"""Benchmark of the DBSCAN hot path on one MI355X.

A step is one full pass of the path over the resident input: classify
(neighbour counts + core flags) and label (union-find + border attachment),
ending with the (cores, noise, clusters) triple on the host; coordinates are
in HBM before the timed region starts and per-point results stay in HBM.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

Prints one JSON line.  `value` is points/s = N / step time.  `roofline` is
computed from HIP events recorded on the library's stream around every
launch of the dominant kernel in the timed steps.  `cpu_baseline` times the
reference's own OpenMP process_points (oracle/_ref/ref_harness) on a bounded
sample of chunk rows of the same input and extrapolates to all N*N pair tests
of its Classify phase; its Assign-Labels phase is not included, so the CPU
figure is an upper bound on the reference's points/s.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nextdbscan-exa_amd"))

METRIC = "points/s clustered, plain vs MFMA-filter/batched, exact labels; % HBM/MFMA peak"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# The 157.3 TFLOP/s vector figure counts packed FP32 (two lanes' worth per
# instruction); this build compiles with packed FP32 off, so one wave64 FP32
# instruction per SIMD per issue slot: 256 CUs x 4 SIMDs x 16 lanes x 2
# flops (FMA) x 2.4 GHz = 78.6 TFLOP/s, i.e. 39.3e12 lane-operations/s.
FP32_NONPACKED_PEAK_TFLOPS = 78.6
BF16_DENSE_PEAK_TFLOPS = 2500.0
DEFAULT_WORKLOAD = "blobs3d_10m"


def algorithmic_bytes(kernel, n, d, cells):
    """Bytes one launch must move at minimum (reads of its inputs, writes of
    its outputs), per kernel: N points of d float32, C_ne non-empty cells."""
    table = {
        "count": n * (4 * d + 4) + 8 * cells,          # sorted coords in, nn out, cell table
        "mf_count": n * (4 * d + 4),                    # coordinates in, nn out
        "union": n * (4 * d + 1) + 8 * cells,          # coords + core flags in, parents
        "border": n * (4 * d + 1 + 4) + 8 * cells,     # coords + core flags in, labels out
        "radix_scatter": n * 24,                        # key+val in and out
        "brute_count": n * (4 * d + 4),
    }
    return table.get(kernel)


def cpu_baseline(x, eps_text, m, n):
    """Reference OpenMP Classify on sampled chunk rows (oracle/_ref)."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        return None
    threads = max(1, min(16, os.cpu_count() or 1))
    rows = max(64, min(n, int(2.5e10 / max(n, 1))))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.bin")
        import synth
        synth.write_bin(p, x)
        out = subprocess.run([harness, "classify_sample", p, str(m), eps_text, str(threads), str(rows), "7"],
                             capture_output=True, text=True, timeout=600, check=True,
                             env=dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores"))
    r = json.loads(out.stdout.strip().splitlines()[-1])
    full_ms = r["ms"] * (n / r["rows"])
    return {
        "value": n / (full_ms * 1e-3),
        "unit": "points/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference process_points (data_process.cpp:395-458, OpenMP, {threads} threads) over "
                   f"{r['rows']} random chunk rows x {n} points = {r['pair_tests']:.3g} pair tests in "
                   f"{r['ms'] / 1e3:.1f} s, extrapolated to the Classify phase's {n}x{n} tests "
                   f"({full_ms / 1e3:.4g} s); Assign Labels (>= one more N x N_core sweep) excluded, "
                   f"so this is an upper bound on the reference's points/s"),
        "pair_tests_per_s": r["tests_per_s"],
    }


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), when a committed
    summary for this kernel and workload exists."""
    p = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def Ctx_stats(ctx):
    import ndbx
    return ndbx.Context.stats(ctx)


def ctx_path(ctx):
    import ndbx
    return ndbx.Context.path(ctx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--n", type=int, default=None, help="smaller instance of the same workload shape")
    ap.add_argument("--n-inst", type=int, default=4096, help="instances of the batched workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU form comparison")
    ap.add_argument("--no-audit", action="store_true", help="skip the float64 audit of the MFMA filter")
    ap.add_argument("--graph", action="store_true",
                    help="also time graph replays (NDBX_GRAPH; 3-D and below) and report them as value")
    ap.add_argument("--autotune", action="store_true",
                    help="batched workload: tile shape from the autotuner (NDBX_AUTOTUNE)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one GPU: every box in this setup has one MI355X, and the multi-GPU "
              "(RCCL/xGMI slab sharding) path is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ndbx
    import synth

    if not ndbx.device_available():
        raise RuntimeError("no HIP device")
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    extra = {}
    if args.workload.startswith("batch3d"):
        # many independent instances: one batched run vs the plain form (each
        # instance its own context), both from device-resident points
        n_inst = args.n_inst
        x, off, eps_arr = synth.batch3d(n_inst=n_inst)
        n, d = x.shape
        m = 8
        mins = np.full(n_inst, m, np.int32)
        tune = ndbx.AUTOTUNE if args.autotune else 0
        ctx = ndbx.Batch(x, off, eps_arr, mins, flags=ndbx.PROFILE | tune)
        step = lambda: (ctx.run(), ctx.meta())
        eps = float(eps_arr.mean())
        eps_text = None
        x_dev = torch.from_numpy(x).to("cuda")
        n_plain = 0 if args.no_plain else n_inst
        plain = [ndbx.Context(None, eps_arr[i], m, device_ptr=x_dev.data_ptr() + int(off[i]) * 4 * d,
                              n=int(off[i + 1] - off[i]), d=d) for i in range(n_plain)]
        for c in plain:  # warm: allocate every context's buffers
            c.classify().label()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 5)):
            for c in plain:
                c.classify().label()
                c.meta()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3 / max(1, args.steps // 5)
        # the plain form must give the batched form's per-instance results
        ctx.run()
        bmeta = ctx.meta()
        for i in range(0, n_plain, max(1, n_inst // 64)):
            assert tuple(bmeta[i]) == plain[i].meta(), f"instance {i} differs"
        for c in plain:
            c.close()
        if n_plain:
            extra["plain_gpu"] = {"ms_per_step": plain_ms, "value": n / (plain_ms * 1e-3),
                                  "form": "one context per instance, run one after another"}
        extra["instances"] = n_inst
        if args.graph:
            # the same batch replayed from a captured graph (first run eager + capture)
            with ndbx.Batch(x, off, eps_arr, mins, flags=ndbx.GRAPH | tune) as gb:
                for _ in range(max(1, args.warmup)):
                    gb.run()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    gb.run()
                    gmeta = gb.meta()
                torch.cuda.synchronize()
                g_ms = (time.perf_counter() - t0) * 1e3 / args.steps
                gc, replays, reruns = gb.info()
                assert np.array_equal(gmeta, bmeta), "graph replay differs from the eager batched run"
            extra["graph"] = {"ms_per_step": g_ms, "value": n / (g_ms * 1e-3), "replays": replays,
                              "reruns": reruns, "group_cells": gc}
    else:
        x, eps, m = synth.make(args.workload, n=args.n)
        n, d = x.shape
        eps_text = {"blobs2d_100k": "0.5", "blobs3d_10m": "1.5", "skew2d_100m": "0.3"}.get(args.workload,
                                                                                             repr(float(eps)))
        ctx = ndbx.Context(x, eps, m, flags=ndbx.PROFILE)
        step = lambda: (ctx.classify().label(), ctx.meta())
        if d > 3 and not args.no_plain:
            # plain GPU form for comparison: the all-pairs float32 kernels
            with ndbx.Context(x, eps, m, flags=ndbx.FORCE_BRUTE) as pc:
                pc.classify().label()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                pc.classify().label()
                pmeta = pc.meta()
                plain_ms = (time.perf_counter() - t0) * 1e3
                plain_nn = pc.nn()
            extra["plain_gpu"] = {"ms_per_step": plain_ms, "value": n / (plain_ms * 1e-3),
                                  "form": "all-pairs float32 kernels (FORCE_BRUTE), 1 timed step"}
        if args.graph and d <= 3:
            # the same pass replayed from a captured graph (one-instance batch:
            # first run eager + capture, then replays)
            off1 = np.array([0, n], np.int64)
            with ndbx.Batch(x, off1, np.array([eps], np.float32), np.array([m], np.int32), flags=ndbx.GRAPH) as gb:
                for _ in range(max(1, args.warmup)):
                    gb.run()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    gb.run()
                    gmeta = gb.meta()
                torch.cuda.synchronize()
                g_ms = (time.perf_counter() - t0) * 1e3 / args.steps
                gc, replays, reruns = gb.info()
                gnn = gb.nn()
            extra["graph"] = {"ms_per_step": g_ms, "value": n / (g_ms * 1e-3), "replays": replays,
                              "reruns": reruns, "group_cells": gc, "meta": [int(v) for v in gmeta[0]]}
            extra["graph_nn"] = gnn
        if d > 3 and not args.no_audit:
            with ndbx.Context(x, eps, m, flags=ndbx.FP64_AUDIT) as ac:
                ac.classify()
                extra["fp64_audit"] = ac.audit()
                extra["fp64_audit"]["scope"] = "every pair of ~2048 sampled 32-row tiles"

    for i in range(args.warmup):
        if i == args.warmup - 1:
            ctx.profile(reset=True)
        step()
    # The timed steps keep HIP events around the dominant kernel only (the
    # counting kernel, else the last warmup step's longest): events around
    # every launch add host work per launch that shows up in the step time.
    # Without warmup every launch is evented.
    warm = ctx.profile(reset=True)
    dominant = next((k for k in ("mf_count", "count") if k in warm), None)
    if dominant is None and warm:
        dominant = max(warm.items(), key=lambda kv: kv[1][0])[0]
    ctx.profile_select(dominant)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, meta = step()
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    prof = ctx.profile(reset=True)
    if dominant is None and prof:
        dominant = next((k for k in ("mf_count", "count") if k in prof), None) or \
            max(prof.items(), key=lambda kv: kv[1][0])[0]
    # per-kernel table: a separate pass of the same step with every launch evented
    ctx.profile_select(None)
    table_steps = max(1, min(args.steps, 3))
    for _ in range(table_steps):
        step()
    ctx.sync()
    table = ctx.profile()
    stats = Ctx_stats(ctx)
    path = ctx_path(ctx)
    nn = ctx.nn()
    if isinstance(meta, np.ndarray):
        meta = [int(v) for v in meta.sum(0)]
    ctx.close()
    if "plain_gpu" in extra and not args.workload.startswith("batch3d"):
        extra["plain_gpu"]["same_result"] = bool(tuple(pmeta) == tuple(meta) and np.array_equal(plain_nn, nn))

    kernels = {k: {"ms_avg": v[0] / max(v[1], 1), "launches_per_step": v[1] / table_steps}
               for k, v in sorted(table.items(), key=lambda kv: -kv[1][0])}
    roofline = None
    if dominant and dominant in prof:
        total_ms, launches = prof[dominant]
        avg_s = total_ms / launches * 1e-3
        ab = algorithmic_bytes(dominant, n, d, stats["cells"])
        if dominant == "mf_count":
            # distance filter: F = 2d flops (a d-term dot product) per unordered
            # point pair of the candidate cell pairs, i.e. d per ordered pair
            # (pair_tests counts ordered pairs); the kernel computes each
            # unordered pair once with 3 bf16 MFMAs per 16-wide k step (hi*hi,
            # hi*lo, lo*hi), k padded to kp: 3 kp issued flops per ordered pair
            kp = 16 if d <= 16 else 32 if d <= 32 else 64 if d <= 64 else 128
            flops = 1.0 * d * float(stats["pair_tests"])
            tf = flops / avg_s / 1e12
            roofline = {"bound": "mfma", "kernel": dominant, "achieved": tf, "peak": BF16_DENSE_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": tf / BF16_DENSE_PEAK_TFLOPS,
                        "traffic": traffic_from_profiles(dominant, args.workload),
                        "algorithmic_flops": flops, "avg_ms": avg_s * 1e3,
                        "issued_mfma_tflops": tf * 3.0 * kp / d,
                        "issued_frac": tf * 3.0 * kp / d / BF16_DENSE_PEAK_TFLOPS}
        elif ab is not None:
            achieved = ab / avg_s / 1e9
            roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": traffic_from_profiles(dominant, args.workload),
                        "algorithmic_bytes": ab, "avg_ms": avg_s * 1e3}
            if dominant in ("count", "brute_count") and stats["pair_tests"]:
                # compute view: the arithmetic issued, d differences, d products
                # and d - 1 sums per pair test (pairs inside dense cells'
                # strip rows are counted from tables, without tests)
                flops = (3.0 * d - 1.0) * float(stats["pair_tests"])
                tf = flops / avg_s / 1e12
                roofline["valu"] = {"achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                    "frac": tf / FP32_PEAK_TFLOPS, "flops": flops,
                                    "basis": "issued pair tests x (3d - 1) flops",
                                    # each flop of a test is its own instruction (no FMA: the
                                    # reference rounds every product and sum), so the ceiling is
                                    # the non-packed lane-operation rate, half the FMA flop rate
                                    "peak_nonpacked_ops": FP32_NONPACKED_PEAK_TFLOPS / 2.0,
                                    "frac_nonpacked_ops": tf / (FP32_NONPACKED_PEAK_TFLOPS / 2.0)}
    cpu = None if (args.no_cpu_baseline or eps_text is None) else cpu_baseline(x, eps_text, m, n)

    gnn = extra.pop("graph_nn", None)
    if gnn is not None:
        # the replays must give the eager pass's results
        assert np.array_equal(gnn, nn) and extra["graph"]["meta"] == list(meta), "graph replay differs"
    config_extra = {}
    if "graph" in extra:
        # value: graph replays; the profiled eager batched run gives the kernel table
        config_extra = {"form": "graph replay", "eager_ms_per_step": ms_per_step,
                        "group_cells": extra["graph"]["group_cells"]}
        ms_per_step = extra["graph"]["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy generator, nextdbscan-exa_amd/synth.py)",
        "config": {"workload": args.workload, "n": n, "d": d, "eps": float(eps), "min_pts": m, "path": path,
                   "cells": stats["cells"], "tiles": stats["tiles"], "pair_tests": stats["pair_tests"], "result": list(meta), **config_extra},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "kernels_pass": f"{table_steps} more steps with every launch evented (the timed steps event only {dominant})",
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
