"""The host CLI (nextdbscan-exa_amd/host/nextdbscan-hip) against the triples the
reference CLI printed for the same inputs (tests/golden/cli_triples.json), and
its -o output against the reference's per-point cluster index."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import synth
from conftest import GOLDEN_DIR, load_golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "nextdbscan-exa_amd", "host", "nextdbscan-hip")
TRIPLES = json.load(open(os.path.join(GOLDEN_DIR, "cli_triples.json")))


def run_cli(args):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)


def test_cli_rejects_missing_parameters():
    # main.cpp:80-84: m and e are required; exit status EXIT_FAILURE
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = run_cli(["-m", "5", "in.bin"])
    assert r.returncode == 1
    assert "Please specify the m and e parameters" in r.stdout


def test_cli_rejects_hdf5():
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = run_cli(["-m", "5", "-e", "0.5", "points.h5"])
    assert r.returncode == 255
    assert "HDF5 is not supported" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["blobs2d_3k", "blobs3d_4k", "lattice2d_exact_eps", "flt_max_padding",
                                  "fp_sens_fma_d3", "blobs_d5", "skew2d_20k", "single_point"])
def test_cli_matches_reference_triple(gpu_lib, tmp_path, name):
    g = load_golden(name)
    p = str(tmp_path / "in.bin")
    synth.write_bin(p, g["x"])
    out = str(tmp_path / "labels.txt")
    r = run_cli(["-m", str(g["m"]), "-e", g["eps_text"], "-t", "4", "-o", out, p])
    assert r.returncode == 0, r.stderr
    cl = int(re.search(r"Estimated Clusters: (-?\d+)", r.stdout).group(1))
    co = int(re.search(r"Core Points: (-?\d+)", r.stdout).group(1))
    no = int(re.search(r"Noise Points: (-?\d+)", r.stdout).group(1))
    assert [cl, co, no] == TRIPLES[name]
    for phase in ["Read Input Data: ", "Classify Points: ", "Assign Labels: ", "Collect Results: ",
                  "Total Execution Time (without I/O): "]:
        assert phase in r.stdout
    labels = np.loadtxt(out, dtype=np.int64).reshape(-1)
    np.testing.assert_array_equal(labels, g["cci"])


@pytest.mark.gpu
def test_cli_csv_input(gpu_lib, tmp_path):
    g = load_golden("blobs2d_3k")
    p = str(tmp_path / "in.csv")
    # %.9g round-trips float32 through atof -> float (magma_input.h:138)
    np.savetxt(p, g["x"], fmt="%.9g", delimiter=" ")
    r = run_cli(["-m", str(g["m"]), "-e", g["eps_text"], p])
    assert r.returncode == 0, r.stderr
    assert "WARNING: USING SLOW CSV I/O." in r.stdout
    co = int(re.search(r"Core Points: (-?\d+)", r.stdout).group(1))
    assert co == TRIPLES["blobs2d_3k"][1]


APPROX = json.load(open(os.path.join(GOLDEN_DIR, "approx_triples.json")))


@pytest.mark.parametrize("name", sorted(APPROX))
def test_reference_approx_contract_holds_on_fixtures(name):
    """The reference's own -a 1 runs (data_process with approx and the CLI):
    cores and noise equal the exact run's, the cluster count is at least the
    exact one (main.cpp:45)."""
    g = load_golden(name)
    cores, noise, clusters = APPROX[name]["data_process_meta"]
    assert (cores, noise) == g["meta"][:2]
    assert clusters >= g["meta"][2]
    cl, co, no = APPROX[name]["cli"]
    assert [co, no] == TRIPLES[name][1:] and cl >= TRIPLES[name][0]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(APPROX))
def test_cli_approx_mode(gpu_lib, tmp_path, name):
    """-a 1 prints the reference's approximate line; cores and noise equal the
    reference's -a 1 run, and the count printed (the exact one) is at most
    the reference's upper bound."""
    g = load_golden(name)
    p = str(tmp_path / "in.bin")
    synth.write_bin(p, g["x"])
    r = run_cli(["-m", str(g["m"]), "-e", g["eps_text"], "-t", "4", "-a", "1", p])
    assert r.returncode == 0, r.stderr
    assert "Estimated Clusters:" not in r.stdout
    cl = int(re.search(r"Upper-bound Approximate Clusters: (-?\d+)", r.stdout).group(1))
    co = int(re.search(r"Core Points: (-?\d+)", r.stdout).group(1))
    no = int(re.search(r"Noise Points: (-?\d+)", r.stdout).group(1))
    ref_cl, ref_co, ref_no = APPROX[name]["cli"]
    assert (co, no) == (ref_co, ref_no)
    assert cl == TRIPLES[name][0] and cl <= ref_cl


@pytest.mark.gpu
def test_approx_flag_same_labels(gpu_lib):
    g = load_golden("blobs3d_4k")
    eps = synth.eps_f32(g["eps_text"])
    with gpu_lib.Context(g["x"], eps, g["m"], flags=gpu_lib.APPROX) as c:
        c.classify().label()
        np.testing.assert_array_equal(c.nn(), g["nn"])
        np.testing.assert_array_equal(c.labels(), g["cci"])
        assert c.meta() == g["meta"]


MEMBERS = os.path.join(ROOT, "tests", "host", "members_check")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["blobs2d_3k", "blobs3d_4k", "duplicates2d", "all_noise", "flt_max_padding"])
def test_host_data_process_members(gpu_lib, tmp_path, name):
    """The host data_process's public members after the reference's call
    sequence: v_coord kept, v_coord_id = 0..n-1, v_coord_nn and
    v_coord_cluster_index equal to the reference's, and v_cluster_label a
    label table over core ranks whose roots are the clusters
    (data_process.cpp:79-88)."""
    g = load_golden(name)
    p, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    synth.write_bin(p, g["x"])
    r = subprocess.run([MEMBERS, p, str(g["m"]), g["eps_text"], out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.int32)
    n, d, n_coord_vals, cores, noise, clusters, n_lab, roots = (int(v) for v in raw[:8])
    o = 8
    ids, nn, cci = raw[o:o + n], raw[o + n:o + 2 * n], raw[o + 2 * n:o + 3 * n]
    lab = raw[o + 3 * n:o + 3 * n + n_lab]
    coords = raw[o + 3 * n + n_lab:].view(np.float32).reshape(n, d)
    assert n_coord_vals == n * d
    np.testing.assert_array_equal(coords, g["x"])
    np.testing.assert_array_equal(ids, np.arange(n))
    np.testing.assert_array_equal(nn, g["nn"])
    np.testing.assert_array_equal(cci, g["cci"])
    core = g["nn"] >= g["m"]
    assert n_lab == core.sum() == cores
    np.testing.assert_array_equal(lab, g["cci"][core])
    assert roots == clusters == g["meta"][2] and (cores, noise) == g["meta"][:2]


@pytest.mark.gpu
def test_cli_chunked_run(gpu_lib, tmp_path):
    """NDBX_CHUNK_POINTS runs the CLI slab by slab: the same printed triple
    and per-point labels as the reference."""
    g = load_golden("skew2d_20k")
    p, out = str(tmp_path / "in.bin"), str(tmp_path / "labels.txt")
    synth.write_bin(p, g["x"])
    r = subprocess.run([CLI, "-m", str(g["m"]), "-e", g["eps_text"], "-o", out, p], capture_output=True, text=True,
                       timeout=120, env=dict(os.environ, NDBX_CHUNK_POINTS="6000"))
    assert r.returncode == 0, r.stderr
    cl = int(re.search(r"Estimated Clusters: (-?\d+)", r.stdout).group(1))
    co = int(re.search(r"Core Points: (-?\d+)", r.stdout).group(1))
    no = int(re.search(r"Noise Points: (-?\d+)", r.stdout).group(1))
    assert [cl, co, no] == TRIPLES["skew2d_20k"]
    np.testing.assert_array_equal(np.loadtxt(out, dtype=np.int64).reshape(-1), g["cci"])
