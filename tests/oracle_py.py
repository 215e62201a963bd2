"""ctypes access to the oracle (oracle/liboracle.so) and to the reference build
under oracle/_ref/.  Test infrastructure: only tests, smoke() and bench.py's
CPU-baseline leg use it."""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_HARNESS = os.path.join(ORACLE_DIR, "_ref", "ref_harness")
REF_CLI = os.path.join(ORACLE_DIR, "_ref", "nextdbscan-omp")
NO_CLUSTER = 2**31 - 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-C", ORACLE_DIR, "oracle"],
                                  stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.oracle_e2.restype = ctypes.c_float
        L.oracle_e2.argtypes = [ctypes.c_float]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.oracle_classify_rows.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                           i64p, ctypes.c_int64, i32p]
        L.oracle_core_label_range_rows.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                                   u8p, i32p, i64p, ctypes.c_int64, i32p, i32p]
        L.oracle_classify.argtypes = [f32p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_float, i32p]
        L.oracle_labels.restype = ctypes.c_int32
        L.oracle_labels.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int32, i32p, i32p]
        L.oracle_dbscan.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int32, i32p, i32p, i32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def e2(eps) -> np.float32:
    return np.float32(lib().oracle_e2(float(np.float32(eps))))


def classify_rows(x, eps, rows):
    """nn of the listed rows only (each against every point)."""
    x = np.ascontiguousarray(x, np.float32)
    rows = np.ascontiguousarray(rows, np.int64)
    nn = np.zeros(len(rows), np.int32)
    lib().oracle_classify_rows(_p(x, ctypes.c_float), x.shape[0], x.shape[1], float(e2(eps)),
                               _p(rows, ctypes.c_int64), len(rows), _p(nn, ctypes.c_int32))
    return nn


def core_label_range_rows(x, eps, core, labels, rows):
    """(smallest, largest) label over the cores within eps of each listed row;
    (INT32_MAX, INT32_MIN) where there is none."""
    x = np.ascontiguousarray(x, np.float32)
    core = np.ascontiguousarray(core, np.uint8)
    labels = np.ascontiguousarray(labels, np.int32)
    rows = np.ascontiguousarray(rows, np.int64)
    lo = np.zeros(len(rows), np.int32)
    hi = np.zeros(len(rows), np.int32)
    lib().oracle_core_label_range_rows(_p(x, ctypes.c_float), x.shape[0], x.shape[1], float(e2(eps)),
                                       _p(core, ctypes.c_uint8), _p(labels, ctypes.c_int32),
                                       _p(rows, ctypes.c_int64), len(rows), _p(lo, ctypes.c_int32),
                                       _p(hi, ctypes.c_int32))
    return lo, hi


def classify(x, eps):
    x = np.ascontiguousarray(x, np.float32)
    nn = np.zeros(x.shape[0], np.int32)
    lib().oracle_classify(_p(x, ctypes.c_float), x.shape[0], x.shape[1],
                          float(e2(eps)), _p(nn, ctypes.c_int32))
    return nn


def dbscan(x, eps, m):
    """Returns (nn, labels, (cores, noise, clusters))."""
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    d = x.shape[1] if x.ndim == 2 else 1
    nn = np.zeros(n, np.int32)
    lab = np.zeros(n, np.int32)
    meta = np.zeros(3, np.int32)
    lib().oracle_dbscan(_p(x, ctypes.c_float), n, d, float(np.float32(eps)), int(m),
                        _p(nn, ctypes.c_int32), _p(lab, ctypes.c_int32),
                        _p(meta, ctypes.c_int32))
    return nn, lab, tuple(int(v) for v in meta)


def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


def ref_full(x, eps_text: str, m: int, threads: int = 8, approx: bool = False):
    """Run the reference's own data_process (oracle/_ref/ref_harness).
    Returns (nn, cci, labels_table, (cores, noise, clusters))."""
    x = np.ascontiguousarray(x, np.float32)
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.bin")
        out = os.path.join(td, "out.bin")
        with open(inp, "wb") as f:
            np.array(x.shape, np.int32).tofile(f)
            x.tofile(f)
        subprocess.check_call([REF_HARNESS, "full", inp, str(m), eps_text, str(threads),
                               "1" if approx else "0", out])
        raw = np.fromfile(out, np.int32)
    n, cores, noise, clusters, nl = (int(v) for v in raw[:5])
    nn = raw[5:5 + n].copy()
    cci = raw[5 + n:5 + 2 * n].copy()
    lab = raw[5 + 2 * n:5 + 2 * n + nl].copy()
    return nn, cci, lab, (cores, noise, clusters)
