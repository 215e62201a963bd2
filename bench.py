"""Benchmark of the DBSCAN hot path on one MI355X.

A step is one full pass of the path over the resident input: classify
(neighbour counts + core flags) and label (union-find + border attachment),
ending with the (cores, noise, clusters) triple on the host; coordinates are
in HBM before the timed region starts and per-point results stay in HBM.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--headline-only]
                    [--no-cpu-baseline]

Prints one JSON line.  `value` is points/s = N / step time on the headline
workload (default: 10M points, d = 3).  `roofline` is computed from HIP
events recorded on the library's stream around every launch of the dominant
kernel in the timed steps; the count kernel is bound by FP32 vector issue
(each pair test is separately rounded sub/mul/add, no FMA), so its binding
roof is the VALU one, with algorithmic flops 3d per ordered neighbour pair
(3d * sum(nn)); its HBM fraction and issued-test arithmetic are reported
beside it.  `cpu_baseline` times the reference's own OpenMP build
(oracle/_ref) on the box's host cores, on a bounded sample of the same
workload.  `forms` holds the other workloads, each with the form it is
compared against: the 100k 2-D blobs (eager and graph replay), the 1M 64-D
clusters (MFMA filter vs the all-pairs float32 kernels, with the float64
audit), the 4,096-instance batch (batched eager and graph replay vs one
context per instance) and the skewed 100M 2-D set.
"""
from __future__ import annotations

import argparse
import json
import os
import re
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
EPS_TEXT = {"blobs2d_100k": "0.5", "blobs3d_10m": "1.5", "skew2d_100m": "0.3"}
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "nextdbscan-omp")


# ------------------------------------------------------------ host facts --
def host_cpu():
    """Cores this process may use, and the host's socket/core layout."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for key, name in (("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
                          ("Thread(s) per core", "threads_per_core"), ("Model name", "model")):
            m = re.search(rf"^{re.escape(key)}:\s*(.+)$", out, re.M)
            if m:
                v = m.group(1).strip()
                info[name] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def physical_cores(info):
    """Physical cores available to this process (affinity / threads per core)."""
    return max(1, info["affinity"] // max(1, int(info.get("threads_per_core", 1) or 1)))


def omp_env(threads):
    return dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")


def write_bin(path, x):
    import synth
    synth.write_bin(path, x)


# ---------------------------------------------------- reference CPU legs --
def cpu_classify_sample(x, eps_text, m, rows, threads, work_note):
    """The reference's process_points (data_process.cpp:395-458) over `rows`
    random chunk rows, extrapolated to its N x N Classify phase."""
    if not os.path.exists(REF_HARNESS):
        return None
    n = len(x)
    rows = max(16, min(n, int(rows)))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.bin")
        write_bin(p, x)
        out = subprocess.run([REF_HARNESS, "classify_sample", p, str(m), eps_text, str(threads), str(rows), "7"],
                             capture_output=True, text=True, timeout=600, check=True, env=omp_env(threads))
    r = json.loads(out.stdout.strip().splitlines()[-1])
    full_ms = r["ms"] * (n / r["rows"])
    return {
        "value": n / (full_ms * 1e-3), "unit": "points/s", "cores": threads, "kind": "reference",
        "sample": (f"reference process_points (OpenMP, {threads} threads) over {r['rows']} random chunk rows x "
                   f"{n} points = {r['pair_tests']:.3g} pair tests in {r['ms'] / 1e3:.1f} s, extrapolated to the "
                   f"Classify phase's {n} x {n} tests ({full_ms / 1e3:.4g} s); {work_note}"),
        "measured": "lower bound: Classify phase only, extrapolated from sampled rows",
        "pair_tests_per_s": r["tests_per_s"],
    }


def cpu_reference_run(x, eps_text, m, threads):
    """The reference CLI (OpenMP build) on the whole input: its own printed
    phase times (nextdbscan.cpp:39-65)."""
    if not os.path.exists(REF_CLI):
        return None
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.bin")
        write_bin(p, x)
        out = subprocess.run([REF_CLI, "-m", str(m), "-e", eps_text, "-t", str(threads), p], capture_output=True,
                             text=True, timeout=900, check=True, env=omp_env(threads)).stdout

    def ms(label):
        mm = re.search(re.escape(label) + r"\s*(\d+)", out)
        return int(mm.group(1)) if mm else None

    tri = [int(re.search(r + r":\s*(-?\d+)", out).group(1))
           for r in ("Core Points", "Noise Points", "(?:Estimated|Upper-bound Approximate) Clusters")]
    return {"total_ms": ms("Total Execution Time (without I/O):"), "classify_ms": ms("Classify Points:"),
            "label_ms": ms("Assign Labels:"), "meta": tri}


# ------------------------------------------------------------- GPU legs --
def algorithmic_bytes(kernel, n, d, cells):
    """Bytes one launch must move at minimum (reads of its inputs, writes of
    its outputs), per kernel: N points of d float32, C_ne non-empty cells."""
    table = {
        "count": n * (4 * d + 4) + 8 * cells,          # sorted coords in, nn out, cell table
        "mf_count": n * (4 * d + 4),                    # coordinates in, nn out
        "brute_count": n * (4 * d + 4),
    }
    return table.get(kernel)


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), when a committed
    summary for this kernel and workload exists."""
    p = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def roofline_for(kernel, avg_s, n, d, stats, sum_nn, workload):
    """Roofline of the dominant kernel from its average launch time."""
    if kernel == "mf_count":
        # F = 2d flops (a d-term dot product) per point pair of the 32 x 32
        # blocks the count pass computes (pair_tests: rows x columns of those
        # blocks, each unordered pair of cells once); blocks the centroid
        # bounds prove empty, or wholly within eps (counted without products),
        # are not computed.  Issued: up to 3 bf16 MFMAs per 16-wide k step
        # (hi*hi, hi*lo, lo*hi), k padded to kp.
        kp = 16 if d <= 16 else 32 if d <= 32 else 64 if d <= 64 else 128
        flops = 2.0 * d * float(stats["pair_tests"])
        tf = flops / avg_s / 1e12
        all_pairs = n * (n - 1) / 2.0
        return {"bound": "mfma", "kernel": kernel, "achieved": tf, "peak": BF16_DENSE_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": tf / BF16_DENSE_PEAK_TFLOPS,
                "traffic": traffic_from_profiles(kernel, workload), "algorithmic_flops": flops,
                "basis": "2d flops per point pair of the blocks the count pass computes", "avg_ms": avg_s * 1e3,
                "computed_pairs": int(stats["pair_tests"]),
                "computed_fraction_of_all_pairs": float(stats["pair_tests"]) / all_pairs,
                "all_pairs_tflops": 2.0 * d * all_pairs / avg_s / 1e12,
                "all_pairs_note": "2d flops per unordered pair of all N(N-1)/2 pairs over the launch: the rate at "
                                  "which pairs are decided, not arithmetic issued",
                "issued_mfma_tflops_max": tf * 3.0 * kp / d,
                "issued_note": "upper bound: every block's three bf16 products issued"}
    ab = algorithmic_bytes(kernel, n, d, stats["cells"])
    if ab is None:
        return None
    gbs = ab / avg_s / 1e9
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "algorithmic_bytes": ab}
    if kernel not in ("count", "brute_count") or sum_nn is None:
        return {"bound": "hbm", "kernel": kernel, "avg_ms": avg_s * 1e3,
                "traffic": traffic_from_profiles(kernel, workload), **hbm}
    # algorithmic arithmetic: d subtractions, d products and d sums per
    # ordered neighbour pair (self included), 3d * sum(nn)
    flops = 3.0 * d * float(sum_nn)
    tf = flops / avg_s / 1e12
    out = {"bound": "valu", "kernel": kernel, "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": tf / FP32_PEAK_TFLOPS, "traffic": traffic_from_profiles(kernel, workload),
           "algorithmic_flops": flops, "basis": "3d flops per ordered neighbour pair, 3d * sum(nn)",
           "sum_nn": int(sum_nn), "avg_ms": avg_s * 1e3,
           # each flop of a test is its own instruction (no FMA: the reference
           # rounds every product and sum), so the ceiling is the non-packed
           # lane-operation rate, half the FMA flop rate
           "peak_nonpacked_ops": FP32_NONPACKED_PEAK_TFLOPS / 2.0,
           "frac_nonpacked_ops": tf / (FP32_NONPACKED_PEAK_TFLOPS / 2.0),
           "hbm": hbm}
    if stats["pair_tests"]:
        # issued: the pair tests the kernel ran (parts of dense cells counted
        # from tables run none), d - 1 sums per test after the first square
        iflops = (3.0 * d - 1.0) * float(stats["pair_tests"])
        out["issued"] = {"pair_tests": int(stats["pair_tests"]), "flops": iflops,
                         "achieved": iflops / avg_s / 1e12,
                         "frac": iflops / avg_s / 1e12 / FP32_PEAK_TFLOPS,
                         # each test decides one unordered pair of distinct points
                         "tests_per_unordered_pair": float(stats["pair_tests"]) / max(1.0, (float(sum_nn) - n) / 2)}
        if out["frac"] > 1.0:
            # dense 2-D cells: strip tables count whole rows of neighbours with
            # no arithmetic, so 3d * sum(nn) is not work the kernel has to do;
            # the arithmetic it issues is the roof's basis instead
            out.update({"achieved": out["issued"]["achieved"], "frac": out["issued"]["frac"],
                        "algorithmic_flops": iflops,
                        "basis": "(3d - 1) flops per issued pair test: dense-cell tables count most neighbour "
                                 "pairs without arithmetic, so 3d * sum(nn) exceeds the work done",
                        "frac_nonpacked_ops": out["issued"]["achieved"] / (FP32_NONPACKED_PEAK_TFLOPS / 2.0)})
    return out


def measure(ctx, step, steps, warmup, table_steps=3):
    """Warm up, time `steps` steps with events around the dominant kernel
    only, then a few more steps with every launch evented (per-kernel table
    and pair-test statistics)."""
    import torch
    for i in range(warmup):
        if i == warmup - 1:
            ctx.profile(reset=True)
        step()
    warm = ctx.profile(reset=True)
    dominant = next((k for k in ("mf_count", "count", "brute_count") if k in warm), None)
    if dominant is None and warm:
        dominant = max(warm.items(), key=lambda kv: kv[1][0])[0]
    ctx.profile_select(dominant)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        meta = step()
    ctx.sync()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    prof = ctx.profile(reset=True)
    ctx.profile_select(None)
    for _ in range(table_steps):
        step()
    ctx.sync()
    table = ctx.profile()
    kernels = {k: {"ms_avg": v[0] / max(v[1], 1), "launches_per_step": v[1] / table_steps}
               for k, v in sorted(table.items(), key=lambda kv: -kv[1][0])}
    avg_s = prof[dominant][0] / prof[dominant][1] * 1e-3 if dominant in prof else None
    return {"ms_per_step": ms, "meta": meta, "dominant": dominant, "avg_s": avg_s, "kernels": kernels,
            "kernels_pass": f"{table_steps} more steps with every launch evented (the timed steps event only "
                            f"{dominant})"}


def bench_single(workload, steps, warmup, n=None, cpu=True, compare_plain=False, audit=False, graph=False,
                 threads=16, top_kernels=None):
    import numpy as np
    import ndbx
    import synth

    x, eps, m = synth.make(workload, n=n)
    n, d = x.shape
    eps_text = EPS_TEXT.get(workload, repr(float(eps)))
    out = {"workload": workload, "n": n, "d": d, "eps": float(eps), "min_pts": m}
    with ndbx.Context(x, eps, m, flags=ndbx.PROFILE) as ctx:
        r = measure(ctx, lambda: ctx.classify().label().meta(), steps, warmup)
        stats = ctx.stats()
        nn = ctx.nn()
        out["path"] = ctx.path()
    sum_nn = int(nn.astype(np.int64).sum())
    meta = list(r["meta"])
    out.update({"ms_per_step": r["ms_per_step"], "value": n / (r["ms_per_step"] * 1e-3), "result": meta,
                "cells": stats["cells"], "tiles": stats["tiles"], "pair_tests": stats["pair_tests"],
                "sum_nn": sum_nn})
    out["roofline"] = roofline_for(r["dominant"], r["avg_s"], n, d, stats, sum_nn, workload) if r["avg_s"] else None
    kernels = r["kernels"]
    if top_kernels:
        kernels = dict(list(kernels.items())[:top_kernels])
    out["kernels"] = kernels
    out["kernels_pass"] = r["kernels_pass"]
    if compare_plain:
        # plain GPU form: the all-pairs float32 kernels, one timed step
        import torch
        with ndbx.Context(x, eps, m, flags=ndbx.FORCE_BRUTE) as pc:
            pc.classify().label()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pc.classify().label()
            pmeta = pc.meta()
            plain_ms = (time.perf_counter() - t0) * 1e3
            same = bool(list(pmeta) == meta and np.array_equal(pc.nn(), nn))
        out["plain_gpu"] = {"ms_per_step": plain_ms, "value": n / (plain_ms * 1e-3), "same_result": same,
                            "form": "all-pairs float32 kernels (FORCE_BRUTE), 1 timed step",
                            "speedup": plain_ms / r["ms_per_step"]}
        assert same, "MFMA filter and all-pairs kernels disagree"
    if audit:
        with ndbx.Context(x, eps, m, flags=ndbx.FP64_AUDIT) as ac:
            ac.classify()
            a = ac.audit()
        stride, offset = os.environ.get("NDBX_AUDIT_STRIDE"), int(os.environ.get("NDBX_AUDIT_OFFSET", "0"))
        a["scope"] = ("every pair of ~512 sampled 128-row groups of the candidate cell pairs" if not stride else
                      f"every pair of the 128-row groups {offset} + {int(stride)} j of the candidate cell pairs")
        a["band"] = "the count pass's: R from the rows' and the column tile's largest centred norm"
        out["fp64_audit"] = a
        assert a["filter_flips"] == 0, f"filter flips: {a}"
    if graph and d <= 3:
        # the same pass replayed from a captured graph (one-instance batch:
        # first run eager + capture, then replays)
        import torch
        off1 = np.array([0, n], np.int64)
        with ndbx.Batch(x, off1, np.array([eps], np.float32), np.array([m], np.int32), flags=ndbx.GRAPH) as gb:
            for _ in range(max(2, warmup)):
                gb.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                gb.run()
                gmeta = gb.meta()
            torch.cuda.synchronize()
            g_ms = (time.perf_counter() - t0) * 1e3 / steps
            _, replays, reruns = gb.info()
            same = bool(np.array_equal(gb.nn(), nn) and [int(v) for v in gmeta[0]] == meta)
        out["graph"] = {"ms_per_step": g_ms, "value": n / (g_ms * 1e-3), "replays": replays, "reruns": reruns,
                        "same_result": same}
        assert same, "graph replay differs from the eager run"
    if cpu:
        if workload == "blobs2d_100k":
            t0 = time.perf_counter()
            ref = cpu_reference_run(x, eps_text, m, threads)
            if ref:
                out["cpu_baseline"] = {
                    "value": n / (ref["total_ms"] * 1e-3), "unit": "points/s", "cores": threads,
                    "kind": "reference", "measured": "full run",
                    "sample": (f"reference CLI (OpenMP build, -t {threads}) on all {n} points: Total Execution "
                               f"Time (without I/O) {ref['total_ms']} ms (Classify {ref['classify_ms']} ms, "
                               f"Assign Labels {ref['label_ms']} ms); wall {time.perf_counter() - t0:.1f} s"),
                    "same_result": ref["meta"] == [meta[0], meta[1], meta[2]],
                }
        else:
            # about 1e11 pair tests of CPU work for d = 2, 3 (~10 s on 16
            # cores); fewer rows for high d, whose tests cost ~d each
            budget = 1.0e11 if d <= 3 else 3.0e11 / d
            out["cpu_baseline"] = cpu_classify_sample(
                x, eps_text, m, budget / n, threads,
                "Assign Labels (at least one more N x N_core sweep) excluded, so this is an upper bound on the "
                "reference's points/s")
    return out


def bench_batch(n_inst, steps, warmup, cpu=True, threads=16, other_forms=True):
    import numpy as np
    import torch
    import ndbx
    import synth

    x, off, eps_arr = synth.batch3d(n_inst=n_inst)
    n, d = x.shape
    m = 8
    mins = np.full(n_inst, m, np.int32)
    out = {"workload": f"batch3d_{n_inst}", "instances": n_inst, "n": n, "d": d, "min_pts": m,
           "eps_range": [float(eps_arr.min()), float(eps_arr.max())]}
    with ndbx.Batch(x, off, eps_arr, mins, flags=ndbx.PROFILE) as b:
        r = measure(b, lambda: b.run().meta(), steps, warmup, table_steps=2)
        bmeta = b.meta()
        nn = b.nn()
        stats = ndbx.Context.stats(b)
    sum_nn = int(nn.astype(np.int64).sum())
    out.update({"eager": {"ms_per_step": r["ms_per_step"], "value": n / (r["ms_per_step"] * 1e-3)},
                "cells": stats["cells"], "tiles": stats["tiles"], "pair_tests": stats["pair_tests"],
                "sum_nn": sum_nn, "result_sum": [int(v) for v in bmeta.sum(0)]})
    out["roofline"] = roofline_for(r["dominant"], r["avg_s"], n, d, stats, sum_nn,
                                   f"batch3d_{n_inst}") if r["avg_s"] else None
    out["kernels"] = dict(list(r["kernels"].items())[:8])
    if not other_forms:  # the eager batched run alone (kernel profiles)
        return out
    # the batch replayed from a captured graph, tile shape from the autotuner
    with ndbx.Batch(x, off, eps_arr, mins, flags=ndbx.GRAPH | ndbx.AUTOTUNE) as gb:
        for _ in range(max(2, warmup)):
            gb.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            gb.run()
            gmeta = gb.meta()
        torch.cuda.synchronize()
        g_ms = (time.perf_counter() - t0) * 1e3 / steps
        gc, replays, reruns = gb.info()
        tw = gb.tile_shape()[1]
        same = bool(np.array_equal(gmeta, bmeta) and np.array_equal(gb.nn(), nn))
    out["graph"] = {"ms_per_step": g_ms, "value": n / (g_ms * 1e-3), "replays": replays, "reruns": reruns,
                    "group_cells": gc, "tiles_per_wave": tw, "same_result": same,
                    "form": "graph replay, autotuned tile shape"}
    assert same, "graph replay differs from the eager batched run"
    # plain form: one context per instance, device-resident points
    x_dev = torch.from_numpy(x).to("cuda")
    plain = [ndbx.Context(None, eps_arr[i], m, device_ptr=x_dev.data_ptr() + int(off[i]) * 4 * d,
                          n=int(off[i + 1] - off[i]), d=d) for i in range(n_inst)]
    for c in plain:  # warm: every context's buffers allocated
        c.classify().label()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pmeta = [c.classify().label().meta() for c in plain]
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    for c in plain:
        c.close()
    del x_dev
    mism = [i for i in range(n_inst) if tuple(bmeta[i]) != tuple(pmeta[i])]
    out["plain_gpu"] = {"ms_per_step": plain_ms, "value": n / (plain_ms * 1e-3),
                        "form": "one context per instance, run one after another, 1 timed step",
                        "same_result": not mism, "instances_compared": n_inst}
    assert not mism, f"batched and per-instance runs differ on instances {mism[:10]}"
    if cpu and os.path.exists(REF_CLI):
        # the reference CLI on whole instances, smallest-first over a spread
        # of sizes until ~15 s of CPU work; extrapolated by sum(n^2), the
        # quadratic cost of both of its phases
        sizes = np.diff(off).astype(np.float64)
        order = np.argsort(sizes)
        pick = order[np.linspace(0, n_inst - 1, 48).round().astype(np.int64)]
        done, tot_ms, sq = [], 0.0, 0.0
        t0 = time.perf_counter()
        for i in pick:
            if time.perf_counter() - t0 > 15.0:
                break
            s, e = int(off[i]), int(off[i + 1])
            ref = cpu_reference_run(x[s:e], repr(float(eps_arr[i])), m, threads)
            tot_ms += ref["total_ms"]
            sq += sizes[i] ** 2
            done.append(int(i))
        full_ms = tot_ms * float((sizes ** 2).sum()) / sq
        out["cpu_baseline"] = {
            "value": n / (full_ms * 1e-3), "unit": "points/s", "cores": threads, "kind": "reference",
            "measured": "sampled instances, extrapolated by sum(n^2)",
            "sample": (f"reference CLI (OpenMP build, -t {threads}) on {len(done)} whole instances (sizes "
                       f"{int(sizes[done].min())}-{int(sizes[done].max())}), sum of Total Execution Time (without "
                       f"I/O) {tot_ms / 1e3:.2f} s, scaled by sum(n^2) of all instances / of the sample = "
                       f"{float((sizes ** 2).sum()) / sq:.1f} to {full_ms / 1e3:.4g} s")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--n", type=int, default=None, help="smaller instance of the headline workload's shape")
    ap.add_argument("--n-inst", type=int, default=4096, help="instances of the batched workload")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads' forms")
    ap.add_argument("--forms", default="blobs2d_100k,clusters64d_1m,batch3d,skew2d_100m",
                    help="comma-separated list of the other workloads to measure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the all-pairs comparison of the MFMA workload, and the graph and per-instance forms "
                         "of the batch")
    ap.add_argument("--no-audit", action="store_true", help="skip the float64 audit of the MFMA filter")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one GPU: every box in this setup has one MI355X, and the multi-GPU "
              "(RCCL/xGMI slab sharding) path is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    import torch  # before the library: both then use the one HIP runtime loaded first
    import ndbx

    if not ndbx.device_available():
        raise RuntimeError("no HIP device")
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    host = host_cpu()
    threads = physical_cores(host)
    cpu = not args.no_cpu_baseline
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[bench {time.perf_counter() - t_start:6.1f}s] {msg}", file=sys.stderr, flush=True)

    progress(f"headline {args.workload}")
    if args.workload.startswith("batch3d"):
        h = bench_batch(args.n_inst, args.steps, args.warmup, cpu=cpu, threads=threads, other_forms=not args.no_plain)
        ms_per_step = h["eager"]["ms_per_step"]
    else:
        h = bench_single(args.workload, args.steps, args.warmup, n=args.n, cpu=cpu, threads=threads,
                         compare_plain=args.workload == "clusters64d_1m" and not args.no_plain,
                         audit=args.workload == "clusters64d_1m" and not args.no_audit)
        ms_per_step = h["ms_per_step"]
    forms = {}
    if not args.headline_only:
        for f in [s for s in args.forms.split(",") if s]:
            if f == args.workload or (f == "batch3d" and args.workload.startswith("batch3d")):
                continue
            progress(f"form {f}")
            if f == "blobs2d_100k":
                forms[f] = bench_single(f, 20, 3, cpu=cpu, graph=True, threads=threads, top_kernels=6)
            elif f == "clusters64d_1m":
                forms[f] = bench_single(f, 5, 2, cpu=cpu, compare_plain=True, audit=True, threads=threads,
                                        top_kernels=6)
            elif f == "batch3d":
                forms[f"batch3d_{args.n_inst}"] = bench_batch(args.n_inst, 5, 2, cpu=cpu, threads=threads)
            elif f == "skew2d_100m":
                forms[f] = bench_single(f, 3, 1, cpu=cpu, threads=threads, top_kernels=6)
            torch.cuda.empty_cache()
    progress("done")
    config = {k: h[k] for k in ("workload", "n", "d", "eps", "min_pts", "path", "cells", "tiles", "pair_tests",
                                "sum_nn", "result", "instances") if k in h}
    out = {
        "metric": METRIC,
        "value": h["n"] / (ms_per_step * 1e-3),
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
        "config": config,
        "roofline": h.get("roofline"),
        "cpu_baseline": h.get("cpu_baseline"),
        "host": host,
        "kernels": h.get("kernels"),
        "kernels_pass": h.get("kernels_pass"),
        "forms": forms,
        "bench_wall_s": time.perf_counter() - t_start,
    }
    for k in ("plain_gpu", "fp64_audit", "graph", "eager"):
        if k in h:
            out[k] = h[k]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
