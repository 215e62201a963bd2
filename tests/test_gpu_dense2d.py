"""Dense 2-D cells (thousands of points per grid cell): the count kernel takes
whole fine rows of such cells without tests when every point of a tile's box
is certainly within eps of them, and tests the rest with the float32
predicate.  nn, labels and meta must stay bit-exact against the oracle."""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu


def run(ndbx, x, eps, m):
    with ndbx.Context(x, eps, m, ndbx.PROFILE) as c:
        c.classify().label()
        return c.nn(), c.labels(), c.meta(), c.path(), c.profile()


def check(ndbx, x, eps, m):
    nn, lab, meta, path, prof = run(ndbx, x, eps, m)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, m)
    assert path == "grid"
    assert "dense_prepare" in prof, sorted(prof)
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_tight_blob_with_halo(gpu_lib):
    # 70k points in one blob of sigma 0.1 (tens of thousands per cell at the
    # centre), a sparser blob and uniform noise around it
    rng = np.random.default_rng(71)
    x = np.concatenate([
        rng.normal(5.0, 0.1, (70000, 2)),
        rng.normal((5.6, 5.2), 0.4, (20000, 2)),
        rng.uniform(3.0, 7.0, (10000, 2)),
    ]).astype(np.float32)
    x = x[rng.permutation(len(x))]
    print("case dense2d tight blob n=100000", flush=True)
    check(gpu_lib, x, synth.eps_f32("0.3"), 200)


def test_lattice_ties(gpu_lib):
    # points on a lattice of step 2^-10, each twice, with eps = 85 steps:
    # squared distances are exact in float32, so thousands of pairs sit at
    # exactly eps^2 (85^2 = 13^2 + 84^2 = 36^2 + 77^2 = 40^2 + 75^2 = 51^2 + 68^2)
    # and must be accepted; cells hold about 7,200 points
    u = 2.0 ** -10
    g = np.arange(260, dtype=np.float32) * np.float32(u) + np.float32(3.0)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    lat = np.stack([xx.ravel(), yy.ravel()], 1)
    rng = np.random.default_rng(72)
    x = np.concatenate([lat, lat[rng.permutation(len(lat))]]).astype(np.float32)
    x = x[rng.permutation(len(x))]
    eps = np.float32(85 * u)
    assert float(eps) * float(eps) == float(np.float32(eps) * np.float32(eps))
    print("case dense2d lattice n=%d" % len(x), flush=True)
    check(gpu_lib, x, eps, 30000)


def test_skewed_sample(gpu_lib):
    # the skewed 100M workload's recipe at 120k points, eps scaled down so the
    # big blob's cells stay dense
    x = synth.skew2d_100m(120000, seed=73)
    x = ((x - 500.0) * np.float32(0.05) + np.float32(500.0)).astype(np.float32)
    print("case dense2d skew n=120000", flush=True)
    check(gpu_lib, x, synth.eps_f32("0.3"), 50)
