import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "nextdbscan-exa_amd"))

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_names():
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False)
    return dict(x=z["x"], eps_text=str(z["eps_text"]), m=int(z["m"]), nn=z["nn"],
                cci=z["cci"], meta=tuple(int(v) for v in z["meta"]))


@pytest.fixture(scope="session")
def gpu_lib():
    """The product library on a HIP device; skips when no device is present."""
    import ndbx
    if not ndbx.device_available():
        pytest.skip("no HIP device")
    return ndbx
