"""Host logic of the library checked without a device: the size limits the
pipelines enforce (csrc/ndbx_limits.h), compiled with g++ into a small
checker program."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nextdbscan-exa_amd", "csrc")
HOST = os.path.join(ROOT, "tests", "host")


def build_and_run(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", CSRC,
                           os.path.join(HOST, name + ".cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_mfma_pair_limit(tmp_path):
    out = build_and_run(tmp_path, "limits_check")
    assert "FAIL" not in out and out.count("ok ") >= 8


def test_slab_plan(tmp_path):
    out = build_and_run(tmp_path, "slabs_check")
    assert "FAIL" not in out and "ok 300 random plans" in out
