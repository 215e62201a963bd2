"""The MFMA distance filter (4 <= d <= 128): every neighbour decision equals
the reference's float32 predicate, so nn, core flags, labels and meta are
bit-exact against the oracle; the float64 audit must report zero decisions
of the bf16 estimate outside its guard band that the float32 predicate
contradicts."""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

CASES = [
    # (d, n, n_blobs, extent, sigma range, noise, eps, m, seed)
    (4, 6000, 8, 40.0, (0.5, 2.0), 0.05, 2.5, 8, 1),
    (8, 8000, 10, 40.0, (0.5, 1.5), 0.05, 3.0, 10, 2),
    (16, 8000, 12, 40.0, (0.5, 1.5), 0.1, 5.0, 10, 3),
    (33, 5000, 6, 30.0, (0.5, 1.0), 0.1, 6.0, 6, 4),
    (64, 6000, 10, 30.0, (0.7, 1.2), 0.1, 10.5, 15, 5),
    (100, 3000, 5, 30.0, (0.5, 1.0), 0.1, 11.0, 5, 6),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"d{c[0]}_n{c[1]}")
def test_mfma_against_oracle(gpu_lib, case):
    d, n, nb, ext, sr, noise, eps, m, seed = case
    x = synth.blobs(n, d, nb, 0.0, ext, None, noise, 900 + seed, sigma_range=sr)
    eps = np.float32(eps)
    print(f"case mfma d={d} n={n}", flush=True)
    with gpu_lib.Context(x, eps, m, flags=gpu_lib.FP64_AUDIT) as c:
        c.classify().label()
        nn, lab, meta, path, audit = c.nn(), c.labels(), c.meta(), c.path(), c.audit()
    nn_o, lab_o, meta_o = O.dbscan(x, eps, m)
    assert path == "mfma-filter"
    assert audit["filter_flips"] == 0, audit
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_mfma_clusters64d(gpu_lib):
    x, eps = synth.clusters64d_1m(20000, seed=1003)
    print("case mfma clusters64d n=20000", flush=True)
    with gpu_lib.Context(x, eps, 15, flags=gpu_lib.FP64_AUDIT) as c:
        c.classify().label()
        nn, lab, meta, audit = c.nn(), c.labels(), c.meta(), c.audit()
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 15)
    assert audit["filter_flips"] == 0, audit
    assert audit["max_abs_err"] < 0.01 * float(eps) ** 2, audit
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_mfma_matches_all_pairs(gpu_lib):
    x = synth.blobs(7000, 24, 9, 0.0, 30.0, None, 0.1, 31, sigma_range=(0.5, 1.5))
    eps = np.float32(6.0)
    a = gpu_lib.dbscan(x, eps, 9)
    b = gpu_lib.dbscan(x, eps, 9, flags=gpu_lib.FORCE_BRUTE)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]


@pytest.mark.parametrize("cap", ["16", "100000000"])
def test_mfma_guard_band_queue(gpu_lib, monkeypatch, cap):
    """A 4-D integer lattice with eps = 1: every axis pair sits exactly at
    eps^2, inside the guard band, so the count pass queues them; a queue of 16
    entries decides the rest in the count pass itself (and blocks with more
    than a wave's queue), a large one defers them all."""
    g = np.arange(12, dtype=np.float32)
    x = np.stack(np.meshgrid(g, g, g, g, indexing="ij"), -1).reshape(-1, 4).copy()
    x = x[np.random.default_rng(7).permutation(len(x))]
    eps = np.float32(1.0)
    monkeypatch.setenv("NDBX_BAND_QUEUE", cap)
    print(f"case mfma lattice4d band queue {cap}", flush=True)
    with gpu_lib.Context(x, eps, 6) as c:
        c.classify().label()
        nn, lab, meta, path = c.nn(), c.labels(), c.meta(), c.path()
        c.classify().label()  # a second run reuses the queue
        assert np.array_equal(c.nn(), nn)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 6)
    assert path == "mfma-filter"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_mfma_union_matches_all_pairs_at_scale(gpu_lib):
    """The 64-D clustered workload at 60K points: counts, labels and the
    cluster count of the filter path equal the all-pairs path's (an
    independent union over points)."""
    x, eps = synth.clusters64d_1m(60000, seed=1003)
    print("case mfma clusters64d n=60000 vs all-pairs", flush=True)
    a = gpu_lib.dbscan(x, eps, 15)
    b = gpu_lib.dbscan(x, eps, 15, flags=gpu_lib.FORCE_BRUTE)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_mfma_union_repeatable(gpu_lib):
    """The union's second launch skips cell pairs whose cores already share a
    set, a state other workgroups change while it runs; every wave of a
    workgroup must take the same decision, or one LDS block mixes two column
    blocks and makes a false union.  Several runs of the 60K-point 64-D case
    must all give the all-pairs labelling."""
    x, eps = synth.clusters64d_1m(60000, seed=1003)
    ref = gpu_lib.dbscan(x, eps, 15, flags=gpu_lib.FORCE_BRUTE)
    for rep in range(8):
        print(f"case mfma clusters64d n=60000 run {rep}", flush=True)
        a = gpu_lib.dbscan(x, eps, 15)
        np.testing.assert_array_equal(a[0], ref[0])
        np.testing.assert_array_equal(a[1], ref[1])
        assert a[2] == ref[2]


def _pruning_mix(seed, d, n_per, eps, far):
    """Clusters of a few widths (sigma in eps units) at centres `far` apart
    (in eps units, along random directions), plus uniform background: tight
    clusters hold tiles wholly within eps, far ones tiles with no pair
    within eps."""
    rng = np.random.default_rng(seed)
    parts = []
    sigmas = (0.02, 0.05, 0.3, 0.7)
    for k, sg in enumerate(sigmas * 2):
        c = rng.normal(0.0, 1.0, size=d)
        c *= far * eps * (k + 1) / np.linalg.norm(c)
        parts.append(rng.normal(0.0, sg * eps / np.sqrt(d), size=(n_per, d)) + c)
    bg = rng.uniform(-far * eps * 9, far * eps * 9, size=(n_per, d))
    x = np.concatenate(parts + [bg]).astype(np.float32)
    return x[rng.permutation(len(x))]


@pytest.mark.parametrize("d,far", [(16, 1.7), (48, 1.05)])
def test_mfma_block_pruning_against_oracle(gpu_lib, d, far):
    """The count pass skips column tiles that the centroid bounds prove hold
    no pair within eps, and counts without products those that hold only
    neighbours: every count, label and the meta equal the oracle's, and the
    pairs computed are far fewer than all pairs."""
    eps = np.float32(4.0)
    x = _pruning_mix(77 + d, d, 1500, float(eps), far)
    n = len(x)
    print(f"case mfma pruning d={d} far={far} n={n}", flush=True)
    with gpu_lib.Context(x, eps, 12, flags=gpu_lib.FP64_AUDIT) as c:
        c.classify().label()
        nn, lab, meta, audit, st = c.nn(), c.labels(), c.meta(), c.audit(), c.stats()
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 12)
    assert audit["filter_flips"] == 0, audit
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o
    assert 0 < st["pair_tests"] < 0.5 * n * n / 2, st


def test_mfma_pruning_near_eps(gpu_lib):
    """Three clusters whose centres are eps apart: pairs across clusters sit
    on both sides of eps, so tiles next to the pruning and the wholly-within
    bounds meet pairs near eps; all counts equal the oracle's."""
    rng = np.random.default_rng(5)
    d, eps = 24, np.float32(3.0)
    cen = np.zeros((3, d))
    cen[1, 0] = float(eps)
    cen[2, 1] = float(eps)
    x = np.concatenate([rng.normal(0.0, 0.15, size=(3000, d)) / np.sqrt(d) * 4 + c for c in cen]).astype(np.float32)
    x = x[rng.permutation(len(x))]
    print("case mfma pruning near eps", flush=True)
    with gpu_lib.Context(x, eps, 10) as c:
        c.classify().label()
        nn, lab, meta = c.nn(), c.labels(), c.meta()
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 10)
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_mfma_label_twice(gpu_lib):
    """A second label call returns the labels of the last classify (the MFMA
    label pass consumes the classify's tile tables), and a new classify and
    label give the same result again."""
    x, eps = synth.clusters64d_1m(12000, seed=1003)
    print("case mfma label twice", flush=True)
    with gpu_lib.Context(x, eps, 15) as c:
        c.classify().label()
        a = (c.nn(), c.labels(), c.meta())
        c.label()
        b = (c.nn(), c.labels(), c.meta())
        c.classify().label()
        e = (c.nn(), c.labels(), c.meta())
    for r in (b, e):
        np.testing.assert_array_equal(a[0], r[0])
        np.testing.assert_array_equal(a[1], r[1])
        assert a[2] == r[2]
