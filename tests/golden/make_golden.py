"""Writes the golden fixtures under tests/golden/ from the reference's own build.

Each fixture is data only: an input point set, the DBSCAN parameters and what
the reference's OpenMP data_process (compiled from its sources by
oracle/Makefile into oracle/_ref/ref_harness) leaves in v_coord_nn and
v_coord_cluster_index, plus its (cores, noise, clusters) triple.  CLI triples
come from the reference CLI built the same way (oracle/_ref/nextdbscan-omp).

Run in a container that has the reference sources and `make -C oracle`:
    python tests/golden/make_golden.py [--approx-only]
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "nextdbscan-exa_amd"))

import oracle_py as O  # noqa: E402
import synth  # noqa: E402

FLT_MAX = np.finfo(np.float32).max


def lattice2d(k=40, spacing=0.5):
    g = np.arange(k, dtype=np.float32) * np.float32(spacing)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    return np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float32)


def lattice3d(k=14, spacing=1.0):
    g = np.arange(k, dtype=np.float32) * np.float32(spacing)
    a, b, c = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([a.ravel(), b.ravel(), c.ravel()], 1).astype(np.float32)


def near_boundary_pairs(d, eps_text, n_pairs, seed, want):
    """Isolated pairs whose separation is within ~3e-7 relative of eps, kept
    when the float32 no-FMA decision differs from `want` ('f64' or 'fma')."""
    rng = np.random.default_rng(seed)
    eps = synth.eps_f32(eps_text)
    e2 = np.float32(eps * eps)
    kept_a, kept_b = [], []
    while len(kept_a) < n_pairs:
        m = 200_000
        a = rng.uniform(0, 1000, size=(m, d)).astype(np.float32)
        u = rng.normal(size=(m, d))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        r = float(eps) * (1.0 + rng.uniform(-3e-7, 3e-7, size=(m, 1)))
        b = (a.astype(np.float64) + r * u).astype(np.float32)
        diff = (a - b).astype(np.float32)
        sq = (diff * diff).astype(np.float32)
        acc = np.zeros(m, np.float32)
        for k in range(d):
            acc = (acc + sq[:, k]).astype(np.float32)
        sep = acc <= e2
        if want == "f64":
            dd = a.astype(np.float64) - b.astype(np.float64)
            other = (dd * dd).sum(1) <= float(e2)
        else:  # fused multiply-add of each square into the running sum
            acc2 = np.zeros(m, np.float32)
            for k in range(d):
                acc2 = (diff[:, k].astype(np.float64) ** 2 + acc2.astype(np.float64)).astype(np.float32)
            other = acc2 <= e2
        idx = np.nonzero(sep != other)[0]
        for i in idx[: n_pairs - len(kept_a)]:
            kept_a.append(a[i])
            kept_b.append(b[i])
    a = np.array(kept_a, np.float32)
    b = np.array(kept_b, np.float32)
    # pairs stay where they were drawn (a translation would re-round the
    # differences); keep only pairs farther than 4*eps from every other pair
    keep = np.ones(len(a), bool)
    for i in range(len(a)):
        dd = np.linalg.norm(a.astype(np.float64) - a[i].astype(np.float64), axis=1)
        dd[i] = np.inf
        if (dd[:i][keep[:i]] < 4.0 * float(eps)).any():
            keep[i] = False
    pts = np.stack([a[keep], b[keep]], 1).reshape(-1, d)
    return np.ascontiguousarray(pts, np.float32)


def cases():
    out = []
    out.append(("blobs2d_3k", synth.blobs2d_100k(3000, seed=11), "0.5", 10))
    x = synth.blobs(4000, 3, 12, 0.0, 60.0, None, 0.03, 12, sigma_range=(1.0, 5.0))
    out.append(("blobs3d_4k", x, "1.5", 20))
    out.append(("lattice2d_exact_eps", lattice2d(), "0.5", 5))
    out.append(("lattice3d_exact_eps", lattice3d(), "1.0", 7))
    out.append(("lattice3d_diag", lattice3d(10), "1.7320508", 9))
    rng = np.random.default_rng(13)
    dup = np.repeat(rng.uniform(0, 10, size=(40, 2)).astype(np.float32), 7, axis=0)
    out.append(("duplicates2d", np.concatenate([dup, rng.uniform(0, 10, (60, 2)).astype(np.float32)]), "0.3", 7))
    out.append(("single_point", np.array([[1.0, 2.0]], np.float32), "0.5", 1))
    out.append(("n_less_than_m", rng.uniform(0, 1, (5, 3)).astype(np.float32), "10", 8))
    out.append(("all_noise", (np.arange(200, dtype=np.float32)[:, None] * np.float32([3.0, 0.0])).astype(np.float32), "1", 2))
    x = synth.blobs2d_100k(1500, seed=14)
    x[::97, 0] = FLT_MAX
    x[5::211, 1] = FLT_MAX
    x[7::301, 0] = -FLT_MAX
    out.append(("flt_max_padding", x, "0.5", 6))
    x = synth.blobs2d_100k(800, seed=15)
    x[::50, 1] = np.nan
    x[3::71, 0] = np.inf
    x[4::83, 1] = -np.inf
    out.append(("nonfinite", x, "0.5", 4))
    out.append(("eps_huge", synth.blobs2d_100k(300, seed=16), "1e20", 250))
    out.append(("eps_zero", np.concatenate([dup[:70], rng.uniform(0, 1, (30, 2)).astype(np.float32)]), "0", 3))
    for d, eps_t, seed in [(2, "0.5", 21), (3, "1.5", 22), (8, "2.0", 23), (64, "14.0", 24)]:
        out.append((f"fp_sens_f64_d{d}", near_boundary_pairs(d, eps_t, 60, seed, "f64"), eps_t, 2))
        out.append((f"fp_sens_fma_d{d}", near_boundary_pairs(d, eps_t, 60, seed + 100, "fma"), eps_t, 2))
    for d, eps_t, m, seed in [(1, "0.05", 5, 31), (4, "2.5", 8, 32), (5, "3.0", 8, 33), (16, "8.0", 6, 34)]:
        x = synth.blobs(2500, d, 8, 0.0, 40.0, None, 0.05, seed, sigma_range=(0.5, 2.0))
        out.append((f"blobs_d{d}", x, eps_t, m))
    x, eps = synth.clusters64d_1m(2000, seed=35)
    out.append(("clusters64d_2k", x, repr(float(eps)), 15))
    x = synth.skew2d_100m(20000, seed=36)
    out.append(("skew2d_20k", x, "0.3", 50))
    return out


def ref_cli_triple(x, eps_text, m):
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.bin")
        synth.write_bin(p, x)
        txt = subprocess.check_output([O.REF_CLI, "-m", str(m), "-e", eps_text, "-t", "8", p],
                                      text=True)
    cl = int(re.search(r"Estimated Clusters: (-?\d+)", txt).group(1))
    co = int(re.search(r"Core Points: (-?\d+)", txt).group(1))
    no = int(re.search(r"Noise Points: (-?\d+)", txt).group(1))
    return [cl, co, no]


APPROX_CASES = ["blobs2d_3k", "blobs3d_4k", "lattice2d_exact_eps", "lattice3d_exact_eps", "duplicates2d",
                "blobs_d5", "skew2d_20k", "all_noise"]


def ref_cli_approx(x, eps_text, m):
    """The reference CLI with -a 1: its printed (clusters, cores, noise)."""
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.bin")
        synth.write_bin(p, x)
        txt = subprocess.check_output([O.REF_CLI, "-m", str(m), "-e", eps_text, "-t", "8", "-a", "1", p],
                                      text=True)
    cl = int(re.search(r"Upper-bound Approximate Clusters: (-?\d+)", txt).group(1))
    co = int(re.search(r"Core Points: (-?\d+)", txt).group(1))
    no = int(re.search(r"Noise Points: (-?\d+)", txt).group(1))
    return [cl, co, no]


def write_approx():
    """approx_triples.json: for each case, the reference data_process run with
    approx = true ((cores, noise, clusters)) and the reference CLI's -a 1
    printed (clusters, cores, noise).  The approximate cluster count depends on
    the sweep's thread schedule; both are one run's values (8 threads)."""
    out = {}
    for name, x, eps_text, m in cases():
        if name not in APPROX_CASES:
            continue
        x = np.ascontiguousarray(x, np.float32)
        _nn, _cci, _lab, meta = O.ref_full(x, eps_text, m, threads=8, approx=True)
        out[name] = {"data_process_meta": [int(v) for v in meta], "cli": ref_cli_approx(x, eps_text, m)}
        print(f"{name:24s} approx meta={meta} cli={out[name]['cli']}", flush=True)
    with open(os.path.join(HERE, "approx_triples.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    assert O.ref_available(), "build oracle/_ref first: make -C oracle"
    if sys.argv[1:] == ["--approx-only"]:
        write_approx()
        return
    triples = {}
    for name, x, eps_text, m in cases():
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim == 1:
            x = x[:, None]
        nn, cci, _lab, meta = O.ref_full(x, eps_text, m, threads=8)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), x=x,
                            eps_text=np.array(eps_text), m=np.int32(m), nn=nn, cci=cci,
                            meta=np.array(meta, np.int32))
        triples[name] = ref_cli_triple(x, eps_text, m)
        print(f"{name:24s} n={x.shape[0]:6d} d={x.shape[1]:3d} eps={eps_text:10s} m={m:3d} "
              f"meta={meta} cli={triples[name]}", flush=True)
    with open(os.path.join(HERE, "cli_triples.json"), "w") as f:
        json.dump(triples, f, indent=1, sort_keys=True)
    write_approx()


if __name__ == "__main__":
    main()
