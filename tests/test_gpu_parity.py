"""Parity of the HIP path against the reference's outputs (golden fixtures) and
the oracle, through the C ABI.  Integer results must be bit-exact: nn, core
flags, noise set, and the cluster index of every point (cores: rank of the
lowest-id core of the component; borders: smallest adjacent cluster index)."""
import numpy as np
import pytest

import oracle_py as O
import synth
from conftest import golden_names, load_golden

pytestmark = pytest.mark.gpu


def run(ndbx, x, eps, m, flags=0):
    with ndbx.Context(x, eps, m, flags) as c:
        c.classify()
        nn_cls = c.nn()
        c.label()
        return nn_cls, c.nn(), c.labels(), c.meta(), c.core(), c.path()


@pytest.mark.parametrize("force_brute", [False, True])
@pytest.mark.parametrize("name", golden_names())
def test_golden_fixture(gpu_lib, name, force_brute):
    g = load_golden(name)
    eps = synth.eps_f32(g["eps_text"])
    flags = gpu_lib.FORCE_BRUTE if force_brute else 0
    print(f"case {name} brute={force_brute} n={g['x'].shape}", flush=True)
    nn0, nn, lab, meta, core, path = run(gpu_lib, g["x"], eps, g["m"], flags)
    np.testing.assert_array_equal(nn0, g["nn"])
    np.testing.assert_array_equal(nn, g["nn"])
    np.testing.assert_array_equal(core, g["nn"] >= g["m"])
    np.testing.assert_array_equal(lab, g["cci"])
    assert meta == g["meta"]
    if force_brute:
        assert path == "all-pairs"


RANDOM_CASES = [
    # (d, n, n_blobs, extent, sigma range, noise, eps text, m, seed)
    (2, 5000, 8, 50.0, (0.5, 2.0), 0.1, "0.5", 10, 1),
    (2, 20000, 30, 100.0, (0.3, 1.0), 0.05, "0.3", 12, 2),
    (3, 8000, 10, 60.0, (1.0, 5.0), 0.02, "1.5", 20, 3),
    (3, 20000, 40, 100.0, (0.5, 3.0), 0.05, "1.0", 8, 4),
    (1, 6000, 10, 100.0, (0.1, 1.0), 0.05, "0.05", 6, 5),
    (3, 3000, 3, 10.0, (0.2, 0.5), 0.0, "0.4", 30, 6),   # dense cells (> 64 points per cell)
    (2, 4000, 2, 5.0, (0.05, 0.1), 0.0, "0.2", 50, 7),   # very dense: hundreds per cell
]


@pytest.mark.parametrize("case", RANDOM_CASES, ids=lambda c: f"d{c[0]}_n{c[1]}_s{c[8]}")
def test_random_against_oracle(gpu_lib, case):
    d, n, nb, ext, sr, noise, eps_t, m, seed = case
    x = synth.blobs(n, d, nb, 0.0, ext, None, noise, 500 + seed, sigma_range=sr)
    eps = synth.eps_f32(eps_t)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, m)
    print(f"case random d={d} n={n} seed={seed}", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, m)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_blobs2d_100k_against_oracle(gpu_lib):
    x, eps, m = synth.make("blobs2d_100k")
    print("case blobs2d_100k", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, m)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, m)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65])
def test_tiny_inputs(gpu_lib, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    eps = np.float32(0.3)
    print(f"case tiny n={n}", flush=True)
    _, nn, lab, meta, _, _ = run(gpu_lib, x, eps, 2)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 2)
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_degenerate_extent_and_duplicates(gpu_lib):
    # every point on one line (zero extent in two dims), many duplicates
    rng = np.random.default_rng(9)
    t = np.round(rng.uniform(0, 20, 3000), 1).astype(np.float32)
    x = np.stack([t, np.zeros_like(t), np.full_like(t, 7.0)], 1)
    eps = np.float32(0.1)
    print("case line-duplicates", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, 5)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 5)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


def test_rerun_is_identical(gpu_lib):
    x, eps, m = synth.make("blobs2d_100k", n=30000)
    with gpu_lib.Context(x, eps, m) as c:
        c.classify().label()
        a = (c.nn(), c.labels(), c.meta())
        c.classify().label()
        b = (c.nn(), c.labels(), c.meta())
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_profile_select(gpu_lib):
    """Events around one kernel name only, then every launch again; the
    results do not depend on what is evented."""
    x, eps, m = synth.make("blobs3d_10m", n=40000)
    with gpu_lib.Context(x, eps, m, flags=gpu_lib.PROFILE) as c:
        c.classify().label()
        every = c.profile(reset=True)
        ref = (c.nn(), c.labels(), c.meta())
        c.profile_select("count").classify().label()
        only = c.profile(reset=True)
        got = (c.nn(), c.labels(), c.meta())
        c.profile_select(None).classify().label()
        again = c.profile()
    assert "count" in every and len(every) > 5
    assert set(only) == {"count"} and only["count"][1] == 1
    assert set(again) == set(every)
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    assert got[2] == ref[2]


def test_state_errors(gpu_lib):
    x = np.zeros((10, 2), np.float32)
    c = gpu_lib.Context(x, 0.5, 2)
    with pytest.raises(gpu_lib.NdbxError) as e:
        c.label()
    assert e.value.code == gpu_lib.ESTATE
    with pytest.raises(gpu_lib.NdbxError):
        c.labels()
    c.classify().label()
    assert c.meta() == (10, 0, 1)
    c.close()


@pytest.mark.parametrize("group_cells", ["1", "3", "16"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_tile_shapes_against_oracle(gpu_lib, monkeypatch, d, group_cells):
    """The symmetric count over tiles of one cell, of 3-cell runs (candidates
    of the tile's own later cells tested per lane) and of 16-cell runs, with
    dense, medium and sparse regions, so tiles of <= 16 and > 16 queries both
    occur."""
    x = synth.blobs(6000, d, 12, 0.0, 40.0, None, 0.1, 700 + d, sigma_range=(0.3, 3.0))
    eps = synth.eps_f32({1: "0.1", 2: "0.6", 3: "1.2"}[d])
    monkeypatch.setenv("NDBX_GROUP_CELLS", group_cells)
    print(f"case tile shapes d={d} group_cells={group_cells}", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, 10)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 10)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


@pytest.mark.parametrize("kind", ["uniform", "lattice"])
def test_dense_cube_parts_against_oracle(gpu_lib, kind):
    """3-D cells of hundreds of points, walked by quarter-cell parts: parts
    wholly within eps of a tile box are counted without tests.  The lattice
    (spacing eps/8, exact in float32) puts points on part boundaries and pairs
    at exactly eps along the axes."""
    eps = synth.eps_f32("1.0")
    if kind == "uniform":
        rng = np.random.default_rng(811)
        x = rng.uniform(0.0, 3.0, size=(30000, 3)).astype(np.float32)
    else:
        g = np.arange(25, dtype=np.float32) * np.float32(0.125)
        x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3).copy()
        x = x[np.random.default_rng(812).permutation(len(x))]
    print(f"case dense cube {kind} n={len(x)}", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, 20)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, 20)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o


@pytest.mark.parametrize("seed", [0, 1])
def test_border_cores_below_and_above(gpu_lib, seed):
    """3-D border points whose only cores lie in the stencil planes below
    their cell (dz < 0: computed only for the border search) or above it,
    in tiles of a few points and in tiles of many, between two dense balls;
    labels must be the oracle's (smallest adjacent cluster id)."""
    rng = np.random.default_rng(40 + seed)
    eps, m = np.float32(1.0), 150
    balls = [np.array(c, np.float32) for c in [(5, 5, 5), (5, 5, 7.6), (9, 5, 5)]]
    parts = [c + rng.normal(0, 0.2, (300, 3)).astype(np.float32) for c in balls]
    # thin shells just above / below the balls: too sparse to hold cores
    for c in balls:
        for dz in (1.0, -1.0):
            sh = rng.normal(0, 0.35, (25, 3)).astype(np.float32)
            sh[:, 2] = c[2] + dz + rng.uniform(-0.1, 0.1, 25).astype(np.float32)
            sh[:, :2] += c[:2]
            parts.append(sh)
    # a dense-ish slab of non-core points between the first two balls
    slab = np.stack([rng.uniform(4.0, 6.0, 60), rng.uniform(4.0, 6.0, 60), rng.uniform(6.2, 6.4, 60)], 1)
    parts.append(slab.astype(np.float32))
    parts.append(rng.uniform(0, 12, (200, 3)).astype(np.float32))
    x = np.ascontiguousarray(np.concatenate(parts, 0))
    print(f"case border below/above seed={seed} n={len(x)}", flush=True)
    _, nn, lab, meta, _, path = run(gpu_lib, x, eps, m)
    nn_o, lab_o, meta_o = O.dbscan(x, eps, m)
    assert path == "grid"
    np.testing.assert_array_equal(nn, nn_o)
    np.testing.assert_array_equal(lab, lab_o)
    assert meta == meta_o
    core = nn_o >= m
    border = (~core) & (lab_o != np.iinfo(np.int32).max)
    assert border.sum() > 50 and core.sum() > 500
