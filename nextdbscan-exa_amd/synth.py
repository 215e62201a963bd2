"""Seeded synthetic workloads for the DBSCAN path, plus the reference's `.bin` format.

The five named workloads follow the sizes, blob shapes and DBSCAN parameters
quoted for this path's headline benchmarks.  Generation is numpy's PCG64 with a
fixed seed per workload, so the same call yields the same points on any host
with this numpy.  Points are shuffled, so input order carries no spatial
locality (as in a file written by an unrelated producer).

`.bin` layout (reference `magma_input.h:113-125`): int32 n, int32 d, then n*d
float32 row-major.
"""
from __future__ import annotations

import numpy as np

# (name, seed) -> parameters.  eps is given as the decimal text the reference
# CLI would receive; it is parsed to float32 exactly as `std::strtof` does
# (main.cpp:71), i.e. round-to-nearest of the decimal value.
WORKLOADS = {
    "blobs2d_100k": dict(n=100_000, d=2, eps="0.5", m=10, seed=1001),
    "blobs3d_10m": dict(n=10_000_000, d=3, eps="1.5", m=20, seed=1002),
    "clusters64d_1m": dict(n=1_000_000, d=64, eps=None, m=15, seed=1003),
    "batch3d_4096": dict(n_inst=4096, d=3, m=8, seed=1004),
    "skew2d_100m": dict(n=100_000_000, d=2, eps="0.3", m=50, seed=1005),
}


def eps_f32(text: str) -> np.float32:
    """float32 of the decimal text via the C library's strtof, the parser the
    reference CLI uses for -e (main.cpp:71)."""
    import ctypes
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    return np.float32(libc.strtof(text.encode(), None))


def e2_of(eps: np.float32) -> np.float32:
    """e2 = e*e in float32 (data_process.h:73)."""
    eps = np.float32(eps)
    return np.float32(eps * eps)


def write_bin(path: str, pts: np.ndarray) -> None:
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n, d = pts.shape
    with open(path, "wb") as f:
        np.array([n, d], dtype=np.int32).tofile(f)
        pts.tofile(f)


def read_bin(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        n, d = np.fromfile(f, dtype=np.int32, count=2)
        pts = np.fromfile(f, dtype=np.float32, count=int(n) * int(d))
    return pts.reshape(int(n), int(d))


def _shuffle(rng, pts):
    perm = rng.permutation(pts.shape[0])
    return np.ascontiguousarray(pts[perm], dtype=np.float32)


def blobs(n, d, n_blobs, lo, hi, sigma, noise_frac, seed, weights=None,
          sigma_range=None):
    """Gaussian blobs with centres U[lo,hi]^d plus uniform noise over the box."""
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * noise_frac))
    n_blob = n - n_noise
    centres = rng.uniform(lo, hi, size=(n_blobs, d))
    if sigma_range is not None:
        sig = rng.uniform(sigma_range[0], sigma_range[1], size=n_blobs)
    else:
        sig = np.full(n_blobs, float(sigma))
    if weights is None:
        counts = np.full(n_blobs, n_blob // n_blobs)
        counts[: n_blob - counts.sum()] += 1
    else:
        counts = rng.multinomial(n_blob, weights)
    parts = []
    for c, s, k in zip(centres, sig, counts):
        parts.append(rng.normal(0.0, s, size=(int(k), d)) + c)
    parts.append(rng.uniform(lo, hi, size=(n_noise, d)))
    pts = np.concatenate(parts, axis=0)
    return _shuffle(rng, pts)


def blobs2d_100k(n=100_000, seed=1001):
    """95% from 10 equal-weight blobs (centres U[0,100]^2, sigma=1), 5% noise."""
    return blobs(n, 2, 10, 0.0, 100.0, 1.0, 0.05, seed)


def blobs3d_10m(n=10_000_000, seed=1002):
    """98% from 200 blobs (centres U[0,1000]^3, sigma~U[1,5], Dirichlet(1)
    weights, multinomial counts), 2% uniform noise."""
    rng = np.random.default_rng(seed + 7)
    w = rng.dirichlet(np.ones(200))
    return blobs(n, 3, 200, 0.0, 1000.0, None, 0.02, seed, weights=w,
                 sigma_range=(1.0, 5.0))


def clusters64d_1m(n=1_000_000, seed=1003, d=64):
    """90% from 50 equal-weight clusters (centres N(0,10^2 I), sigma~U[0.5,1.5]),
    10% N(0, 20^2 I) background.  Returns (points, eps) with eps the 2nd
    percentile (linear interpolation) of the float64 pairwise distances of the
    first 10k shuffled points, cast to float32."""
    rng = np.random.default_rng(seed)
    n_bg = n // 10
    n_cl = n - n_bg
    centres = rng.normal(0.0, 10.0, size=(50, d))
    sig = rng.uniform(0.5, 1.5, size=50)
    counts = np.full(50, n_cl // 50)
    counts[: n_cl - counts.sum()] += 1
    parts = [rng.normal(0.0, s, size=(int(k), d)) + c
             for c, s, k in zip(centres, sig, counts)]
    parts.append(rng.normal(0.0, 20.0, size=(n_bg, d)))
    pts = _shuffle(rng, np.concatenate(parts, axis=0))
    sub = pts[: min(10_000, n)].astype(np.float64)
    sq = (sub * sub).sum(1)
    g = sub @ sub.T
    d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * g, 0.0)
    iu = np.triu_indices(sub.shape[0], 1)
    dist = np.sqrt(d2[iu])
    eps = np.float32(np.percentile(dist, 2.0))
    return pts, eps


def batch3d(n_inst=4096, seed=1004, lo_n=256, hi_n=65536, d=3):
    """Independent instances: sizes log-uniform in [lo_n, hi_n]; per instance
    1-20 blobs in [0,100]^3 with sigma~U[0.5,3], 5% noise, eps~U[0.5,2.0].
    Returns (points, offsets[n_inst+1] int64, eps[n_inst] float32)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo_n), np.log(hi_n), size=n_inst)).astype(np.int64)
    sizes = np.clip(sizes, lo_n, hi_n)
    eps = rng.uniform(0.5, 2.0, size=n_inst).astype(np.float32)
    nb = rng.integers(1, 21, size=n_inst)
    parts = []
    for i in range(n_inst):
        parts.append(blobs(int(sizes[i]), d, int(nb[i]), 0.0, 100.0, None, 0.05,
                           seed * 100003 + i, sigma_range=(0.5, 3.0)))
    offsets = np.zeros(n_inst + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    return np.concatenate(parts, axis=0), offsets, eps


def skew2d_100m(n=100_000_000, seed=1005):
    """50% one blob (sigma=0.5 at (500,500)), 45% over 10,000 blobs (sigma=2,
    centres U[0,1000]^2, Zipf(s=1.2) weights), 5% uniform noise."""
    rng = np.random.default_rng(seed)
    n_big = n // 2
    n_noise = n // 20
    n_small = n - n_big - n_noise
    ranks = np.arange(1, 10_001, dtype=np.float64)
    w = ranks ** -1.2
    w /= w.sum()
    counts = rng.multinomial(n_small, w)
    centres = rng.uniform(0.0, 1000.0, size=(10_000, 2))
    out = np.empty((n, 2), dtype=np.float32)
    out[:n_big] = rng.normal(0.0, 0.5, size=(n_big, 2)) + 500.0
    idx = np.repeat(np.arange(10_000), counts)
    out[n_big:n_big + n_small] = rng.normal(0.0, 2.0, size=(n_small, 2)) + centres[idx]
    out[n_big + n_small:] = rng.uniform(0.0, 1000.0, size=(n_noise, 2))
    perm = rng.permutation(n)
    return out[perm]


def make(name: str, n: int | None = None):
    """Return (points float32[n,d], eps float32, min_pts int) for a workload.

    `n` overrides the point count (a smaller instance of the same shape)."""
    if name == "blobs2d_100k":
        return blobs2d_100k(n or 100_000), eps_f32("0.5"), 10
    if name == "blobs3d_10m":
        return blobs3d_10m(n or 10_000_000), eps_f32("1.5"), 20
    if name == "clusters64d_1m":
        pts, eps = clusters64d_1m(n or 1_000_000)
        return pts, eps, 15
    if name == "skew2d_100m":
        return skew2d_100m(n or 100_000_000), eps_f32("0.3"), 50
    raise KeyError(name)
