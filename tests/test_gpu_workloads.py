"""The benchmark workloads themselves against the oracle.

* 10M-point 3-D blobs: a crop of the real input (8 whole blobs, ~197k points)
  bit-exact against the oracle's full run; the full 10M run checked through
  properties the oracle can decide per point: nn of sampled rows, core =
  nn >= m everywhere, every sampled core's label shared by all cores within
  eps of it, every sampled non-core point's label the smallest among its
  cores (noise when none), and canonical cluster ids (a cluster's id is the
  core rank of its lowest core).
* 4,096-instance 3-D batch (sizes log-uniform 256-65,536, eps in [0.5, 2],
  1-20 blobs): the full batch eager and replayed from a graph with the
  autotuned tile shape, 64+ instances (the 8 largest among them) bit-exact
  against the oracle, and the two runs identical on every point.
"""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

I32_MIN = np.iinfo(np.int32).min


@pytest.fixture(scope="module")
def blobs3d_10m():
    return synth.make("blobs3d_10m")


def run(ndbx, x, eps, m, flags=0):
    with ndbx.Context(x, eps, m, flags) as c:
        c.classify().label()
        return c.nn(), c.labels(), c.meta(), c.core(), c.path()


def test_blobs3d_10m_crop_against_oracle(gpu_lib, blobs3d_10m):
    x, eps, m = blobs3d_10m
    lo = np.array([575.0, 750.0, 175.0], np.float32)
    hi = lo + np.float32(250.0)
    xc = np.ascontiguousarray(x[np.all((x >= lo) & (x < hi), axis=1)])
    assert 150_000 <= len(xc) <= 400_000, len(xc)
    print(f"case blobs3d_10m crop n={len(xc)}", flush=True)
    nn, lab, meta, core, path = run(gpu_lib, xc, eps, m)
    assert path == "grid"
    nn_o, lab_o, meta_o = O.dbscan(xc, eps, m)
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(core, nn_o >= m)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o
    assert meta_o[2] >= 5  # several clusters, not one blob


def check_full_run(x, eps, m, nn, lab, meta, core, seed, n_rows=800):
    """Size-independent checks of a whole run; returns the rows checked."""
    n = len(x)
    np.testing.assert_array_equal(core, nn >= m)
    noise = lab == O.NO_CLUSTER
    assert meta[0] == int(core.sum()) and meta[1] == int(noise.sum())
    assert not noise[core].any()
    # canonical ids: a cluster's id is the core rank of its lowest core
    cl = lab[core]
    ids, first = np.unique(cl, return_index=True)
    np.testing.assert_array_equal(ids, first)
    assert meta[2] == len(ids)
    # sampled rows of each kind: any point, cores, borders, noise
    rng = np.random.default_rng(seed)
    border = np.nonzero(~core & ~noise)[0]
    groups = [rng.choice(n, n_rows, replace=False)]
    for pool in (np.nonzero(core)[0], border, np.nonzero(noise)[0]):
        if len(pool):
            groups.append(rng.choice(pool, min(n_rows, len(pool)), replace=False))
    rows = np.unique(np.concatenate(groups))
    np.testing.assert_array_equal(nn[rows], O.classify_rows(x, eps, rows))
    lo, hi = O.core_label_range_rows(x, eps, core, lab, rows)
    rc = core[rows]
    np.testing.assert_array_equal(lo[rc], lab[rows][rc])
    np.testing.assert_array_equal(hi[rc], lab[rows][rc])
    np.testing.assert_array_equal(lo[~rc], lab[rows][~rc])
    return rows


@pytest.mark.timeout(600)
def test_blobs3d_10m_full_properties(gpu_lib, blobs3d_10m):
    x, eps, m = blobs3d_10m
    print(f"case blobs3d_10m full n={len(x)}", flush=True)
    nn, lab, meta, core, path = run(gpu_lib, x, eps, m)
    assert path == "grid"
    print(f"  meta {meta}; oracle rows", flush=True)
    rows = check_full_run(x, eps, m, nn, lab, meta, core, seed=2024)
    assert len(rows) >= 2000


@pytest.fixture(scope="module")
def batch3d_4096():
    x, off, eps = synth.batch3d(n_inst=4096)
    return x, off, eps, np.full(len(eps), 8, np.int32)


def sampled_instances(off, eps, k=72):
    """The 8 largest instances, the 4 smallest-eps and 4 largest-eps ones of
    more than 20,000 points, and the rest spread evenly over the sizes."""
    sizes = np.diff(off)
    order = np.argsort(sizes, kind="stable")
    largest = order[-8:]
    big = np.nonzero(sizes > 20_000)[0]
    by_eps = big[np.argsort(eps[big], kind="stable")]
    ends = np.concatenate([by_eps[:4], by_eps[-4:]])
    spread = order[np.linspace(0, len(order) - 9, k - 16).round().astype(np.int64)]
    return np.unique(np.concatenate([largest, ends, spread]))


@pytest.mark.timeout(900)
def test_batch3d_4096_against_oracle(gpu_lib, batch3d_4096):
    x, off, eps, m = batch3d_4096
    print(f"case batch3d_4096 n={len(x)} eager", flush=True)
    with gpu_lib.Batch(x, off, eps, m) as b:
        b.run()
        nn, lab, meta = b.nn(), b.labels(), b.meta()
    print("case batch3d_4096 graph+autotune", flush=True)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH | gpu_lib.AUTOTUNE) as g:
        g.run()
        np.testing.assert_array_equal(g.nn(), nn)
        np.testing.assert_array_equal(g.labels(), lab)
        np.testing.assert_array_equal(g.meta(), meta)
        g.run()
        gc, replays, reruns = g.info()
        assert replays == 1 and reruns == 0, (replays, reruns)
        np.testing.assert_array_equal(g.nn(), nn)
        np.testing.assert_array_equal(g.labels(), lab)
        np.testing.assert_array_equal(g.meta(), meta)
    inst = sampled_instances(off, eps)
    assert len(inst) >= 64
    sizes = np.diff(off)
    assert sizes[inst].max() == sizes.max()
    for i in inst:
        s, e = off[i], off[i + 1]
        nn_o, lab_o, meta_o = O.dbscan(x[s:e], eps[i], m[i])
        np.testing.assert_array_equal(nn[s:e], nn_o, err_msg=f"instance {i} n={e - s}")
        np.testing.assert_array_equal(lab[s:e], lab_o, err_msg=f"instance {i} n={e - s}")
        assert tuple(meta[i]) == meta_o, i
    # eps at both ends of its range among the checked instances
    assert eps[inst].min() < 0.7 and eps[inst].max() > 1.8
