"""Batched runs replayed from a captured HIP graph (NDBX_GRAPH) and the tile
autotuner (NDBX_AUTOTUNE): every replay must equal an eager run on the same
coordinates, instance by instance, including after new coordinates, after
counts outgrow the captured capacities, and after an input the captured key
layout cannot hold."""
import numpy as np
import pytest

import oracle_py as O
import synth

pytestmark = pytest.mark.gpu


def eager(ndbx, x, off, eps, m):
    with ndbx.Batch(x, off, eps, m) as b:
        b.run()
        return b.nn(), b.labels(), b.meta()


def check_same(got, want, what):
    np.testing.assert_array_equal(got[0], want[0], err_msg=what + " nn")
    np.testing.assert_array_equal(got[1], want[1], err_msg=what + " labels")
    np.testing.assert_array_equal(got[2], want[2], err_msg=what + " meta")


def batch_inputs(seed, n_inst=60, spread=1.0):
    x, off, eps = synth.batch3d(n_inst=n_inst, seed=seed, lo_n=256, hi_n=6000)
    x = (x * np.float32(spread)).astype(np.float32)
    m = np.full(len(eps), 8, np.int32)
    return x, off, eps, m


def test_graph_replays_equal_eager(gpu_lib):
    x, off, eps, m = batch_inputs(101)
    want = eager(gpu_lib, x, off, eps, m)
    print(f"case graph replay n={len(x)}", flush=True)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH) as b:
        for r in range(4):
            b.run()
            check_same((b.nn(), b.labels(), b.meta()), want, f"run {r}")
            # unrelated eager work on the device between replays
            eager(gpu_lib, x[: off[3]], off[:4], eps[:3], m[:3])
        gc, replays, reruns = b.info()
        assert replays == 3 and reruns == 0, (replays, reruns)


def test_graph_new_coordinates_and_capacity(gpu_lib):
    x, off, eps, m = batch_inputs(102)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH) as b:
        b.run()
        b.run()
        # same instances, new points: replayed from the graph
        rng = np.random.default_rng(103)
        x2 = (x + rng.normal(0.0, 0.3, x.shape)).astype(np.float32)
        print("case graph new coords", flush=True)
        b.set_coords(x2).run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x2, off, eps, m), "new coords")
        # points spread 8x apart: many more cells than captured -> eager rerun
        x3 = (x2 * np.float32(8.0)).astype(np.float32)
        print("case graph over capacity", flush=True)
        b.set_coords(x3).run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x3, off, eps, m), "over capacity")
        _, replays, reruns = b.info()
        assert reruns >= 1, (replays, reruns)
        # and the recaptured graph replays the spread input
        b.run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x3, off, eps, m), "recaptured")


def test_graph_input_outside_layout(gpu_lib):
    # a NaN puts an instance outside the grid path: the replay reports it, the
    # eager rerun refuses the batch as an eager batch does, and valid
    # coordinates afterwards run (and capture) again
    x, off, eps, m = batch_inputs(104, n_inst=12)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH) as b:
        b.run()
        b.run()
        x2 = x.copy()
        x2[off[5] + 1, 0] = np.nan
        print("case graph nan input", flush=True)
        with pytest.raises(gpu_lib.NdbxError) as e_graph:
            b.set_coords(x2).run()
        with pytest.raises(gpu_lib.NdbxError) as e_eager:
            eager(gpu_lib, x2, off, eps, m)
        assert e_graph.value.code == e_eager.value.code == gpu_lib.EUNSUPPORTED
        _, replays, reruns = b.info()
        assert reruns == 1, (replays, reruns)
        b.set_coords(x).run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x, off, eps, m), "after nan")
        b.run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x, off, eps, m), "after nan, replay")
        assert b.info()[1] == replays + 1


def test_autotune_keeps_results(gpu_lib):
    x, off, eps, m = batch_inputs(105)
    want = eager(gpu_lib, x, off, eps, m)
    print("case autotune", flush=True)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.AUTOTUNE | gpu_lib.GRAPH) as b:
        b.run()
        gc, _, _ = b.info()
        assert gc in (1, 2, 4, 8, 16)
        assert b.tile_shape() in [(gc, 1), (gc, 4)]
        check_same((b.nn(), b.labels(), b.meta()), want, "tuned")
        b.run()
        check_same((b.nn(), b.labels(), b.meta()), want, "tuned replay")


def test_graph_2d_batch(gpu_lib):
    rng = np.random.default_rng(106)
    sizes = rng.integers(100, 3000, size=30)
    parts = [synth.blobs(int(k), 2, 3, 0.0, 30.0, None, 0.05, 1060 + i, sigma_range=(0.3, 1.5))
             for i, k in enumerate(sizes)]
    x = np.concatenate(parts).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    eps = rng.uniform(0.3, 1.0, size=30).astype(np.float32)
    m = np.full(30, 6, np.int32)
    want = eager(gpu_lib, x, off, eps, m)
    print("case graph 2d", flush=True)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH) as b:
        for r in range(3):
            b.run()
            check_same((b.nn(), b.labels(), b.meta()), want, f"2d run {r}")
        assert b.info()[1] == 2


def test_graph_over_capacity_dense_cells_past_cap(gpu_lib):
    """A replay whose cell count outgrows the capacity while the cells past
    the capacity are dense (64+ points, so they have quarter-cell part
    tables) must stay in bounds, flag the run and give the eager result.
    Instance 0 spans [0, 200]^3 from the start, so the key layout is wide
    enough for the later spread; instances 1-8 go from [0, 4]^3 (few cells)
    to [0, 160]^3 (one cell per point); the last instance stays a dense cube
    and its cells (numbered last) land past the capacity."""
    rng = np.random.default_rng(107)
    parts = [rng.uniform(0.0, 200.0, size=(500, 3))]
    parts += [rng.uniform(0.0, 4.0, size=(3000, 3)) for _ in range(8)]
    parts.append(rng.uniform(0.0, 2.0, size=(8000, 3)))
    sizes = [len(p) for p in parts]
    x = np.concatenate(parts).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    eps = np.full(len(sizes), 1.0, np.float32)
    m = np.full(len(sizes), 8, np.int32)
    x2 = x.copy()
    x2[off[1]:off[9]] *= np.float32(40.0)
    print("case graph over capacity, dense cells past the cap", flush=True)
    with gpu_lib.Batch(x, off, eps, m, gpu_lib.GRAPH) as b:
        b.run()
        b.run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x, off, eps, m), "compact")
        b.set_coords(x2).run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x2, off, eps, m), "spread")
        _, replays, reruns = b.info()
        assert replays == 1 and reruns == 1, (replays, reruns)
        b.run()
        check_same((b.nn(), b.labels(), b.meta()), eager(gpu_lib, x2, off, eps, m), "spread, recaptured")
