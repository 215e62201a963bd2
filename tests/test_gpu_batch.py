"""Batched form: many independent instances in one run must give, instance by
instance, exactly what the oracle (and a single-instance run) gives."""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu


def make_batch(n_inst, d, seed, max_n=2500):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, max_n, size=n_inst)
    sizes[: min(3, n_inst)] = [0, 1, 2][: min(3, n_inst)]  # empty and tiny instances
    parts, eps, m = [], [], []
    for i, n in enumerate(sizes):
        nb = int(rng.integers(1, 6))
        parts.append(synth.blobs(int(n), d, nb, 0.0, 40.0, None, 0.05, int(seed * 1000 + i),
                                 sigma_range=(0.3, 2.0)) if n else np.zeros((0, d), np.float32))
        eps.append(np.float32(rng.uniform(0.3, 1.5)))
        m.append(int(rng.integers(2, 12)))
    off = np.zeros(n_inst + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return np.concatenate(parts, 0).astype(np.float32), off, np.array(eps, np.float32), np.array(m, np.int32)


@pytest.mark.parametrize("d,seed", [(2, 1), (3, 2), (1, 3)])
def test_batch_matches_oracle_per_instance(gpu_lib, d, seed):
    x, off, eps, m = make_batch(40, d, seed)
    print(f"case batch d={d} seed={seed} n={len(x)}", flush=True)
    with gpu_lib.Batch(x, off, eps, m) as b:
        b.run()
        nn, lab, meta = b.nn(), b.labels(), b.meta()
    for i in range(len(off) - 1):
        s, e = off[i], off[i + 1]
        nn_o, lab_o, meta_o = O.dbscan(x[s:e], eps[i], m[i])
        np.testing.assert_array_equal(nn[s:e], nn_o, err_msg=f"instance {i}")
        np.testing.assert_array_equal(lab[s:e], lab_o, err_msg=f"instance {i}")
        assert tuple(meta[i]) == meta_o, i


def test_batch_matches_single_runs(gpu_lib):
    x, off, eps = synth.batch3d(n_inst=24, seed=77, lo_n=256, hi_n=20000)
    m = np.full(len(eps), 8, np.int32)
    print(f"case batch-vs-single n={len(x)}", flush=True)
    with gpu_lib.Batch(x, off, eps, m) as b:
        b.run()
        nn, lab, meta = b.nn(), b.labels(), b.meta()
        b.run()  # a second run on the same context is identical
        np.testing.assert_array_equal(b.labels(), lab)
    for i in range(len(eps)):
        s, e = off[i], off[i + 1]
        nn1, lab1, meta1 = gpu_lib.dbscan(x[s:e], eps[i], 8)
        np.testing.assert_array_equal(nn[s:e], nn1)
        np.testing.assert_array_equal(lab[s:e], lab1)
        assert tuple(meta[i]) == meta1


def test_batch_rejects_bad_offsets(gpu_lib):
    x = np.zeros((10, 2), np.float32)
    with pytest.raises(gpu_lib.NdbxError) as e:
        gpu_lib.Batch(x, np.array([0, 5, 3, 10]), 0.5, 3)
    assert e.value.code == gpu_lib.EINVAL


@pytest.mark.parametrize("d", [2, 3])
def test_batch_tiny_instances(gpu_lib, d):
    """Hundreds of instances of 0-40 points: most waves of the per-point
    passes hold several instances, most count tiles one or two cells."""
    rng = np.random.default_rng(40 + d)
    sizes = rng.integers(0, 41, size=300)
    parts = [rng.normal(0.0, 1.0, size=(int(k), d)).astype(np.float32) * np.float32(rng.uniform(0.2, 2.0))
             for k in sizes]
    x = np.concatenate(parts, 0)
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    eps = rng.uniform(0.2, 1.0, size=len(sizes)).astype(np.float32)
    m = rng.integers(2, 6, size=len(sizes)).astype(np.int32)
    print(f"case batch tiny d={d} n={len(x)}", flush=True)
    with gpu_lib.Batch(x, off, eps, m) as b:
        b.run()
        nn, lab, meta = b.nn(), b.labels(), b.meta()
    for i in range(len(sizes)):
        s, e = off[i], off[i + 1]
        nn_o, lab_o, meta_o = O.dbscan(x[s:e], eps[i], m[i])
        np.testing.assert_array_equal(nn[s:e], nn_o, err_msg=f"instance {i}")
        np.testing.assert_array_equal(lab[s:e], lab_o, err_msg=f"instance {i}")
        assert tuple(meta[i]) == meta_o, i


@pytest.mark.parametrize("quad", ["0", "1"])
def test_batch_tiles_per_wave_against_oracle(gpu_lib, monkeypatch, quad):
    """Small 3-D tiles counted one per wave and four per wave (16 lanes
    each) give the oracle's counts and labels, instance by instance."""
    x, off, eps, m = make_batch(24, 3, 11, max_n=1500)
    monkeypatch.setenv("NDBX_QUAD_TILES", quad)
    print(f"case batch tiles per wave quad={quad} n={len(x)}", flush=True)
    with gpu_lib.Batch(x, off, eps, m) as b:
        b.run()
        assert b.tile_shape()[1] == (4 if quad == "1" else 1)
        nn, lab, meta = b.nn(), b.labels(), b.meta()
    for i in range(len(off) - 1):
        lo, hi = int(off[i]), int(off[i + 1])
        nn_o, lab_o, meta_o = O.dbscan(x[lo:hi], eps[i], int(m[i]))
        np.testing.assert_array_equal(nn[lo:hi], nn_o)
        np.testing.assert_array_equal(lab[lo:hi], lab_o)
        assert tuple(meta[i]) == meta_o
