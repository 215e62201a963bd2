"""Chunked runs (ndbx_set_chunking): slab by slab, with at least 4 slabs,
the results must be bit-identical to the resident run and to the oracle:
nn, core flags, every point's cluster index and the meta triple."""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu


def run(ndbx, x, eps, m, chunk=0):
    with ndbx.Context(x, eps, m) as c:
        c.set_chunking(chunk)
        c.classify().label()
        return c.nn(), c.labels(), c.meta(), c.core(), c.slabs()


def check_same(got, want, what):
    np.testing.assert_array_equal(got[0], want[0], err_msg=what + " nn")
    np.testing.assert_array_equal(got[1], want[1], err_msg=what + " labels")
    assert got[2] == want[2], what
    np.testing.assert_array_equal(got[3], want[3], err_msg=what + " core")


def test_chunked_blobs3d_10m_crop(gpu_lib):
    x, eps, m = synth.make("blobs3d_10m")
    lo = np.array([575.0, 750.0, 175.0], np.float32)
    xc = np.ascontiguousarray(x[np.all((x >= lo) & (x < lo + np.float32(250.0)), axis=1)])
    print(f"case chunked blobs3d_10m crop n={len(xc)}", flush=True)
    want = run(gpu_lib, xc, eps, m)
    got = run(gpu_lib, xc, eps, m, chunk=60_000)
    assert got[4] >= 4, got[4]
    check_same(got, want, "crop")


def test_chunked_skew2d_sample(gpu_lib):
    """The skewed 2-D shape (one dense blob, many small ones, noise): slabs
    through the dense blob hold little more than its halo."""
    x, eps, m = synth.make("skew2d_100m", n=200_000)
    print(f"case chunked skew2d n={len(x)}", flush=True)
    want = run(gpu_lib, x, eps, m)
    got = run(gpu_lib, x, eps, m, chunk=70_000)
    assert got[4] >= 4, got[4]
    check_same(got, want, "skew2d")


@pytest.mark.parametrize("d,chunk", [(1, 700), (2, 1500), (3, 2500), (3, 900)])
def test_chunked_against_oracle(gpu_lib, d, chunk):
    x = synth.blobs(8000, d, 10, 0.0, 40.0, None, 0.1, 930 + d, sigma_range=(0.3, 2.0))
    eps = synth.eps_f32({1: "0.05", 2: "0.5", 3: "1.0"}[d])
    print(f"case chunked vs oracle d={d} chunk={chunk}", flush=True)
    nn, lab, meta, core, slabs = run(gpu_lib, x, eps, 8, chunk=chunk)
    assert slabs >= 4
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 8)
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_chunked_exact_eps_lattice(gpu_lib):
    """Pairs at exactly eps along x straddle slab edges: spacing 0.5 = eps."""
    g = np.arange(60, dtype=np.float32) * np.float32(0.5)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float32)
    print("case chunked lattice", flush=True)
    got = run(gpu_lib, x, np.float32(0.5), 5, chunk=500)
    assert got[4] >= 4
    nn_o, lab_o, meta_o = O.dbscan(x, np.float32(0.5), 5)
    np.testing.assert_array_equal(got[0], nn_o)
    np.testing.assert_array_equal(got[1], lab_o)
    assert got[2] == meta_o


def test_chunked_rerun_and_errors(gpu_lib):
    x = synth.blobs(5000, 2, 5, 0.0, 30.0, None, 0.05, 940, sigma_range=(0.3, 1.0))
    eps = np.float32(0.4)
    with gpu_lib.Context(x, eps, 6) as c:
        c.set_chunking(1200).classify().label()
        a = (c.nn(), c.labels(), c.meta())
        assert c.slabs() >= 4
        c.classify().label()  # a second chunked run on the same context
        np.testing.assert_array_equal(c.labels(), a[1])
        c.set_chunking(0).classify().label()  # back to resident
        np.testing.assert_array_equal(c.nn(), a[0])
        np.testing.assert_array_equal(c.labels(), a[1])
        assert c.meta() == a[2] and c.slabs() == 0
        # a budget below the points within eps of one x value
        c.set_chunking(3)
        with pytest.raises(gpu_lib.NdbxError) as e:
            c.classify()
        assert e.value.code == gpu_lib.EUNSUPPORTED and "budget" in str(e.value)
    # d > 3 does not take the grid
    y = np.random.default_rng(1).normal(size=(500, 8)).astype(np.float32)
    with gpu_lib.Context(y, 1.0, 4) as c:
        c.set_chunking(100)
        with pytest.raises(gpu_lib.NdbxError) as e:
            c.classify()
        assert e.value.code == gpu_lib.EUNSUPPORTED
