"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, and reports errors without a device (no compute is launched)."""
import ctypes
import os
import re

import pytest

import ndbx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "ndbx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ndbx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ndbx.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(ndbx.SIGNATURES), "binding and header disagree"


def test_version_and_error_string():
    assert ndbx.lib().ndbx_version() >= 100
    assert isinstance(ndbx.lib().ndbx_last_error(), bytes)


def test_argument_errors_are_status_codes():
    L = ndbx.lib()
    h = ctypes.c_void_p()
    assert L.ndbx_create(None, None, 0, 2, 0.5, 5, 0) == ndbx.EINVAL
    assert b"null" in L.ndbx_last_error()
    assert L.ndbx_create(ctypes.byref(h), None, -1, 2, 0.5, 5, 0) == ndbx.EINVAL
    assert L.ndbx_create(ctypes.byref(h), None, 10, 0, 0.5, 5, 0) == ndbx.EINVAL
    assert L.ndbx_create(ctypes.byref(h), None, 10, 2, 0.5, 5, 0) == ndbx.EINVAL  # null coords
    assert L.ndbx_create(ctypes.byref(h), None, 0, 2, 0.5, 5, 1 << 20) == ndbx.EINVAL  # unknown flag bit
    assert b"flag" in L.ndbx_last_error()
    assert L.ndbx_classify(None) == ndbx.EINVAL
    assert L.ndbx_label(None) == ndbx.EINVAL
    assert L.ndbx_destroy(None) == ndbx.OK  # like free(NULL)


@pytest.mark.skipif(ndbx.lib().ndbx_device_count() > 0, reason="a device is present")
def test_no_device_is_reported_not_hidden():
    import numpy as np
    with pytest.raises(ndbx.NdbxError) as e:
        ndbx.Context(np.zeros((4, 2), np.float32), 0.5, 2)
    assert e.value.code == ndbx.EDEVICE
    assert not ndbx.device_available()
