"""The oracle pinned: against the reference's own outputs (golden fixtures),
closed forms, brute force in numpy, and planted mistakes that must fail."""
import numpy as np
import pytest

import oracle_py as O
import synth
from conftest import golden_names, load_golden


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    g = load_golden(name)
    eps = synth.eps_f32(g["eps_text"])
    nn, lab, meta = O.dbscan(g["x"], eps, g["m"])
    np.testing.assert_array_equal(nn, g["nn"])
    np.testing.assert_array_equal(lab, g["cci"])
    assert meta == g["meta"]


def test_e2_is_float_square():
    # data_process.h:73: e2 = e*e in float
    for t in ["0.5", "0.3", "1.5", "1.7320508", "1e20", "0"]:
        e = synth.eps_f32(t)
        assert O.e2(e) == np.float32(e * e)
    assert np.isinf(O.e2(synth.eps_f32("1e20")))


def test_lattice_closed_form_2d():
    # spacing 0.5, eps 0.5: the 4 axis neighbours sit at d^2 = 0.25 = e2 exactly
    # (inclusive '<='), diagonals at 0.5 > e2.  nn = 1 + #axis neighbours.
    k = 30
    g = np.arange(k, dtype=np.float32) * np.float32(0.5)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel()], 1)
    nn = O.classify(x, np.float32(0.5))
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    expect = 1 + (i > 0) + (i < k - 1) + (j > 0) + (j < k - 1)
    np.testing.assert_array_equal(nn, expect.ravel())
    # m=5: interior points are cores, one component; edge points are borders
    nn2, lab, (cores, noise, clusters) = O.dbscan(x, np.float32(0.5), 5)
    assert cores == (k - 2) ** 2 and clusters == 1
    assert noise == 4  # corners touch no core
    assert (lab[nn2 >= 5] == 0).all()


def test_lattice_closed_form_3d_diag():
    # spacing 1, eps = float(sqrt(3)) rounded: all 26 neighbours are within
    # eps iff fl(fl(1)+fl(1)+fl(1)) = 3 <= fl(eps)^2.
    k = 8
    g = np.arange(k, dtype=np.float32)
    a, b, c = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([a.ravel(), b.ravel(), c.ravel()], 1)
    eps = synth.eps_f32("1.7320508")
    e2 = np.float32(eps * eps)
    nn = O.classify(x, eps)
    ii = [np.minimum(v, 1) + np.minimum(k - 1 - v, 1) + 1 for v in (a, b, c)]
    full = (ii[0] * ii[1] * ii[2]).ravel()
    axis_only = (1 + (ii[0] - 1) + (ii[1] - 1) + (ii[2] - 1)).ravel()
    two = None
    if e2 >= np.float32(3.0):
        np.testing.assert_array_equal(nn, full)
    else:
        # diagonals (3) excluded, face-diagonals (2) included
        two = full - ((ii[0] - 1) * (ii[1] - 1) * (ii[2] - 1)).ravel()
        np.testing.assert_array_equal(nn, two)
    assert axis_only.max() == 7


def _numpy_classify(x, e2, mode):
    """Brute force with a chosen arithmetic; mode 'sep' is the reference's."""
    n, d = x.shape
    nn = np.zeros(n, np.int64)
    valid = x[:, 0] != np.finfo(np.float32).max
    for i in range(n):
        diff = (x[i][None, :] - x).astype(np.float32)
        if mode == "f64":
            acc = ((x[i][None, :].astype(np.float64) - x.astype(np.float64)) ** 2).sum(1)
            ok = acc <= float(e2)
        else:
            acc = np.zeros(n, np.float32)
            for k in range(d):
                if mode == "fma":
                    acc = (diff[:, k].astype(np.float64) ** 2 + acc.astype(np.float64)).astype(np.float32)
                else:
                    acc = (acc + (diff[:, k] * diff[:, k]).astype(np.float32)).astype(np.float32)
            ok = acc < e2 if mode == "strict" else acc <= e2
        nn[i] = (ok & valid).sum()
    return nn


@pytest.mark.parametrize("name", ["fp_sens_f64_d3", "fp_sens_fma_d3", "fp_sens_fma_d64",
                                  "lattice2d_exact_eps", "flt_max_padding"])
def test_numpy_restatement_agrees_and_mutants_fail(name):
    """The separately-rounded float32 sum reproduces the reference's nn, and
    each plausible mistake (FMA contraction, float64, strict '<', dropping the
    padding skip) disagrees with it on at least one fixture point."""
    g = load_golden(name)
    x = g["x"]
    e = synth.eps_f32(g["eps_text"])
    e2 = np.float32(e * e)
    good = _numpy_classify(x, e2, "sep")
    np.testing.assert_array_equal(good, g["nn"])
    mutants = {"f64": "fp_sens_f64", "fma": "fp_sens_fma", "strict": "lattice2d"}
    for mode, fixture_prefix in mutants.items():
        if name.startswith(fixture_prefix):
            assert (_numpy_classify(x, e2, mode) != g["nn"]).any(), mode
    if name == "flt_max_padding":
        # without the padding skip, FLT_MAX points would count each other
        n = x.shape[0]
        no_skip = np.zeros(n, np.int64)
        for i in range(n):
            diff = (x[i][None, :] - x).astype(np.float32)
            acc = np.zeros(n, np.float32)
            for k in range(x.shape[1]):
                acc = (acc + (diff[:, k] * diff[:, k]).astype(np.float32)).astype(np.float32)
            no_skip[i] = (acc <= e2).sum()
        assert (no_skip != g["nn"]).any()


def test_labels_against_bfs_components():
    """Core partition = connected components (BFS in numpy), ids = rank of the
    lowest-id core; borders take the smallest adjacent cluster id."""
    x = synth.blobs(1200, 2, 6, 0.0, 30.0, None, 0.1, 77, sigma_range=(0.5, 1.5))
    eps = np.float32(0.6)
    m = 6
    nn, lab, (cores, noise, clusters) = O.dbscan(x, eps, m)
    e2 = np.float32(eps * eps)
    n = x.shape[0]
    diff = (x[:, None, :] - x[None, :, :]).astype(np.float32)
    acc = np.zeros((n, n), np.float32)
    for k in range(2):
        acc = (acc + (diff[..., k] * diff[..., k]).astype(np.float32)).astype(np.float32)
    adj = acc <= e2
    np.testing.assert_array_equal(adj.sum(1), nn)
    core = nn >= m
    rank = np.cumsum(core) - 1
    comp = -np.ones(n, np.int64)
    for s in np.nonzero(core)[0]:
        if comp[s] >= 0:
            continue
        comp[s] = rank[s]
        stack = [s]
        while stack:
            u = stack.pop()
            for v in np.nonzero(adj[u] & core)[0]:
                if comp[v] < 0:
                    comp[v] = rank[s]
                    stack.append(v)
    expect = np.full(n, O.NO_CLUSTER, np.int64)
    expect[core] = comp[core]
    for i in np.nonzero(~core)[0]:
        nb = np.nonzero(adj[i] & core)[0]
        if len(nb):
            expect[i] = comp[nb].min()
    np.testing.assert_array_equal(lab, expect)
    assert clusters == len(np.unique(comp[core]))
    assert cores == core.sum() and noise == (expect == O.NO_CLUSTER).sum()


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_oracle_matches_reference_build_live():
    x = synth.blobs2d_100k(6000, seed=99)
    nn, lab, meta = O.dbscan(x, synth.eps_f32("0.5"), 10)
    rnn, rcci, _, rmeta = O.ref_full(x, "0.5", 10, threads=8)
    np.testing.assert_array_equal(nn, rnn)
    np.testing.assert_array_equal(lab, rcci)
    assert meta == rmeta


@pytest.mark.parametrize("name", golden_names())
def test_classify_rows_matches_reference_fixture(name):
    """The row-subset count equals the reference's nn on every fixture, in
    any row order and with repeated rows."""
    g = load_golden(name)
    eps = synth.eps_f32(g["eps_text"])
    n = len(g["nn"])
    rows = np.random.default_rng(5).permutation(n)
    rows = np.concatenate([rows, rows[: min(n, 3)]])
    np.testing.assert_array_equal(O.classify_rows(g["x"], eps, rows), g["nn"][rows])


@pytest.mark.parametrize("name", golden_names())
def test_core_label_range_rows_matches_reference_fixture(name):
    """On the reference's own labelling: every core's label is both the
    smallest and the largest label among the cores within eps of it, and a
    non-core point's label is the smallest (NO_CLUSTER when there is none)."""
    g = load_golden(name)
    eps = synth.eps_f32(g["eps_text"])
    core = g["nn"] >= g["m"]
    lab = g["cci"]
    rows = np.arange(len(lab))
    lo, hi = O.core_label_range_rows(g["x"], eps, core, lab, rows)
    np.testing.assert_array_equal(lo[core], lab[core])
    np.testing.assert_array_equal(hi[core], lab[core])
    np.testing.assert_array_equal(lo[~core], lab[~core])
    assert (hi[~core & (lab == O.NO_CLUSTER)] == np.iinfo(np.int32).min).all()


def test_core_label_range_rows_sees_planted_mistakes():
    """Checks built on the range function catch a split cluster (one core
    relabelled) and a wrong border (a border given a larger adjacent id)."""
    g = load_golden("blobs2d_3k")
    eps = synth.eps_f32(g["eps_text"])
    core = g["nn"] >= g["m"]
    rows = np.arange(len(core))
    bad = g["cci"].copy()
    victim = np.nonzero(core & (g["nn"] > g["m"] + 2))[0][0]
    bad[victim] = bad[victim] + 1
    lo, hi = O.core_label_range_rows(g["x"], eps, core, bad, rows)
    assert ((lo[core] != bad[core]) | (hi[core] != bad[core])).any()
    # two clusters meeting at a border point: give it the larger id
    lo0, hi0 = O.core_label_range_rows(g["x"], eps, core, g["cci"], rows)
    amb = np.nonzero(~core & (lo0 != hi0) & (hi0 != np.iinfo(np.int32).min))[0]
    if len(amb):
        bad2 = g["cci"].copy()
        bad2[amb[0]] = hi0[amb[0]]
        assert (lo0[~core] != bad2[~core]).any()
