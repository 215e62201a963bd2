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
