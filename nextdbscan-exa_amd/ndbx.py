"""Python binding of libndbx.so (ctypes over the C ABI in include/ndbx.h).

This mirrors the reference's host flow (nextdbscan.cpp:46-59): construct a
context from the points (data_process ctor), classify (select_and_process),
label (select_cores_and_process), then read the meta triple
(get_result_meta) and the per-point arrays.  There is no CPU fallback: if the
library is missing or no HIP device is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libndbx.so")
NO_CLUSTER = 2**31 - 1

OK, EINVAL, ENOMEM, EDEVICE, ESTATE, EUNSUPPORTED = range(6)
APPROX = 1 << 0
FORCE_BRUTE = 1 << 1
DEVICE_INPUT = 1 << 2
PROFILE = 1 << 3
MFMA_FILTER = 1 << 4
FP64_AUDIT = 1 << 5
GRAPH = 1 << 6
AUTOTUNE = 1 << 7
PATH_NAMES = {0: "grid", 1: "all-pairs", 2: "mfma-filter"}

# Every entry point the header declares, with its argument types.
_I64 = ctypes.c_int64
_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_IP = ctypes.POINTER(ctypes.c_int)
_I64P = ctypes.POINTER(ctypes.c_int64)
_DP = ctypes.POINTER(ctypes.c_double)
_CPP = ctypes.POINTER(ctypes.c_char_p)
SIGNATURES = {
    "ndbx_version": ([], ctypes.c_int),
    "ndbx_last_error": ([], ctypes.c_char_p),
    "ndbx_device_count": ([], ctypes.c_int),
    "ndbx_create": ([_PP, _P, _I64, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_uint32], ctypes.c_int),
    "ndbx_classify": ([_P], ctypes.c_int),
    "ndbx_label": ([_P], ctypes.c_int),
    "ndbx_result_meta": ([_P, _IP, _IP, _IP, _IP], ctypes.c_int),
    "ndbx_get_nn": ([_P, _P], ctypes.c_int),
    "ndbx_get_labels": ([_P, _P], ctypes.c_int),
    "ndbx_get_core": ([_P, _P], ctypes.c_int),
    "ndbx_device_results": ([_P, _PP, _PP], ctypes.c_int),
    "ndbx_path": ([_P, _IP], ctypes.c_int),
    "ndbx_sync": ([_P], ctypes.c_int),
    "ndbx_profile_count": ([_P, _IP], ctypes.c_int),
    "ndbx_profile_entry": ([_P, ctypes.c_int, _CPP, _DP, _I64P], ctypes.c_int),
    "ndbx_profile_reset": ([_P], ctypes.c_int),
    "ndbx_profile_select": ([_P, ctypes.c_char_p], ctypes.c_int),
    "ndbx_stats": ([_P, _I64P, _I64P, _I64P], ctypes.c_int),
    "ndbx_audit": ([_P, _I64P, _I64P, _I64P, _DP], ctypes.c_int),
    "ndbx_destroy": ([_P], ctypes.c_int),
    "ndbx_batch_create": ([_PP, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_uint32], ctypes.c_int),
    "ndbx_batch_run": ([_P], ctypes.c_int),
    "ndbx_batch_meta": ([_P, _P], ctypes.c_int),
    "ndbx_batch_set_coords": ([_P, _P], ctypes.c_int),
    "ndbx_batch_info": ([_P, _IP, _I64P, _I64P], ctypes.c_int),
    "ndbx_tile_shape": ([_P, _IP, _IP], ctypes.c_int),
    "ndbx_set_chunking": ([_P, _I64], ctypes.c_int),
    "ndbx_chunk_info": ([_P, _I64P], ctypes.c_int),
    "ndbx_dbscan_noalloc": ([_P, _I64, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_uint32, _P, _P, _P,
                             _P], ctypes.c_int),
}
NOALLOC_MAX_POINTS = 131072


class NdbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ndbx error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libndbx.so (built in-tree by `make -C nextdbscan-exa_amd`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                               "or make -C nextdbscan-exa_amd")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise NdbxError(rc, lib().ndbx_last_error().decode(errors="replace"))


def device_available() -> bool:
    return lib().ndbx_device_count() > 0


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One DBSCAN run on the device (reference: class data_process)."""

    def __init__(self, coords, eps, min_pts, flags=0, device_ptr=None, n=None, d=None):
        self._h = ctypes.c_void_p()
        if device_ptr is not None:
            self.n, self.d = int(n), int(d)
            _check(lib().ndbx_create(ctypes.byref(self._h), ctypes.c_void_p(device_ptr), self.n, self.d,
                                     float(np.float32(eps)), int(min_pts), flags | DEVICE_INPUT))
        else:
            x = np.ascontiguousarray(coords, dtype=np.float32)
            if x.ndim == 1:
                x = x[:, None]
            self.n, self.d = x.shape
            _check(lib().ndbx_create(ctypes.byref(self._h), _ptr(x) if self.n else None, self.n, self.d,
                                     float(np.float32(eps)), int(min_pts), flags))

    def classify(self):
        _check(lib().ndbx_classify(self._h))
        return self

    def label(self):
        _check(lib().ndbx_label(self._h))
        return self

    def meta(self):
        """(cores, noise, clusters), data_process.cpp:50-91."""
        a, b, c, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().ndbx_result_meta(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                      ctypes.byref(n)))
        return a.value, b.value, c.value

    def nn(self):
        out = np.empty(self.n, np.int32)
        _check(lib().ndbx_get_nn(self._h, _ptr(out)))
        return out

    def labels(self):
        out = np.empty(self.n, np.int32)
        _check(lib().ndbx_get_labels(self._h, _ptr(out)))
        return out

    def core(self):
        out = np.empty(self.n, np.uint8)
        _check(lib().ndbx_get_core(self._h, _ptr(out)))
        return out.astype(bool)

    def path(self):
        p = ctypes.c_int()
        _check(lib().ndbx_path(self._h, ctypes.byref(p)))
        return PATH_NAMES.get(p.value, str(p.value))

    def sync(self):
        _check(lib().ndbx_sync(self._h))

    def set_chunking(self, max_slab_points):
        """Slab-by-slab runs with at most max_slab_points points per slab (0: resident)."""
        _check(lib().ndbx_set_chunking(self._h, int(max_slab_points)))
        return self

    def slabs(self):
        k = ctypes.c_int64()
        _check(lib().ndbx_chunk_info(self._h, ctypes.byref(k)))
        return k.value

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().ndbx_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return dict(cells=a.value, tiles=b.value, pair_tests=c.value)

    def audit(self):
        a, b, c, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        _check(lib().ndbx_audit(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(e)))
        return dict(band_pairs=a.value, fp64_flips=b.value, filter_flips=c.value, max_abs_err=e.value)

    def profile(self, reset=False):
        """{kernel name: (total ms, launches)} from HIP events (PROFILE flag)."""
        cnt = ctypes.c_int()
        _check(lib().ndbx_profile_count(self._h, ctypes.byref(cnt)))
        out = {}
        for i in range(cnt.value):
            nm, ms, k = ctypes.c_char_p(), ctypes.c_double(), ctypes.c_int64()
            _check(lib().ndbx_profile_entry(self._h, i, ctypes.byref(nm), ctypes.byref(ms), ctypes.byref(k)))
            out[nm.value.decode()] = (ms.value, k.value)
        if reset:
            _check(lib().ndbx_profile_reset(self._h))
        return out

    def profile_select(self, name=None):
        """Events only around launches named `name` (None: every launch)."""
        _check(lib().ndbx_profile_select(self._h, name.encode() if name else None))
        return self

    def close(self):
        if self._h:
            _check(lib().ndbx_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dbscan(coords, eps, min_pts, flags=0):
    """Whole path on the device: returns (nn, labels, (cores, noise, clusters))."""
    with Context(coords, eps, min_pts, flags) as c:
        c.classify().label()
        return c.nn(), c.labels(), c.meta()


class Batch:
    """Many independent instances in one run (ndbx_batch_*)."""

    def __init__(self, coords, offsets, eps, min_pts, flags=0):
        x = np.ascontiguousarray(coords, dtype=np.float32)
        if x.ndim == 1:
            x = x[:, None]
        self.off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.n_inst = len(self.off) - 1
        self.eps = np.ascontiguousarray(np.broadcast_to(np.float32(eps) if np.isscalar(eps) else eps,
                                                        (self.n_inst,)), dtype=np.float32)
        self.m = np.ascontiguousarray(np.broadcast_to(min_pts, (self.n_inst,)), dtype=np.int32)
        self.n, self.d = x.shape
        self._h = ctypes.c_void_p()
        _check(lib().ndbx_batch_create(ctypes.byref(self._h), _ptr(x) if self.n else None, _ptr(self.off),
                                       self.n_inst, self.d, _ptr(self.eps), _ptr(self.m), flags))

    def run(self):
        _check(lib().ndbx_batch_run(self._h))
        return self

    def meta(self):
        out = np.empty((self.n_inst, 3), np.int32)
        _check(lib().ndbx_batch_meta(self._h, _ptr(out)))
        return out

    def set_coords(self, coords):
        """New coordinates for the same instances (same shape)."""
        x = np.ascontiguousarray(coords, dtype=np.float32).reshape(self.n, self.d)
        _check(lib().ndbx_batch_set_coords(self._h, _ptr(x) if self.n else None))
        return self

    def info(self):
        """(cells per counting group, graph replays, graph replays rerun eagerly)."""
        gc = ctypes.c_int()
        rp = ctypes.c_int64()
        rr = ctypes.c_int64()
        _check(lib().ndbx_batch_info(self._h, ctypes.byref(gc), ctypes.byref(rp), ctypes.byref(rr)))
        return gc.value, rp.value, rr.value

    def tile_shape(self):
        """(cells per counting group, tiles per wave: 1, 4, or 0 while left to the tiles' mean size)."""
        gc = ctypes.c_int()
        tw = ctypes.c_int()
        _check(lib().ndbx_tile_shape(self._h, ctypes.byref(gc), ctypes.byref(tw)))
        return gc.value, tw.value

    def nn(self):
        out = np.empty(self.n, np.int32)
        _check(lib().ndbx_get_nn(self._h, _ptr(out)))
        return out

    def labels(self):
        out = np.empty(self.n, np.int32)
        _check(lib().ndbx_get_labels(self._h, _ptr(out)))
        return out

    def profile(self, reset=False):
        return Context.profile(self, reset)

    def profile_select(self, name=None):
        return Context.profile_select(self, name)

    def sync(self):
        _check(lib().ndbx_sync(self._h))

    def close(self):
        if self._h:
            _check(lib().ndbx_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
