"""ndbx_dbscan_noalloc (no device allocation, fixed scratch array) against the
context entry points on the same device-resident input: nn, labels and the
meta triple bit-exact, on consecutive calls of shrinking size, at the largest
accepted size with every point in its own cell, on the all-pairs kernels, and
with sizes and flags it must refuse."""
import ctypes

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


class Hip:
    """The HIP runtime the library itself loaded (same soname), for the test's
    own device buffers and stream."""

    def __init__(self, ndbx):
        ndbx.lib()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
                  "hipDeviceSynchronize"):
            getattr(self.L, f).restype = ctypes.c_int
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]

    def ok(self, rc):
        assert rc == 0, f"HIP call failed: {rc}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 256)))
        return p.value

    def free(self, p):
        self.ok(self.L.hipFree(ctypes.c_void_p(p)))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.ok(self.L.hipMemcpy(ctypes.c_void_p(p), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1))
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype)
        if n:
            self.ok(self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), out.nbytes, 2))
        return out

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.L.hipStreamCreate(ctypes.byref(s)))
        return s.value


@pytest.fixture(scope="module")
def hip(gpu_lib):
    return Hip(gpu_lib)


def noalloc(ndbx, hip, x_dev, n, d, eps, m, flags=0, stream=None):
    nn = hip.alloc(4 * n)
    lab = hip.alloc(4 * n)
    meta = np.zeros(3, np.int32)
    try:
        rc = ndbx.lib().ndbx_dbscan_noalloc(ctypes.c_void_p(x_dev), n, d, float(np.float32(eps)), int(m), flags,
                                            ctypes.c_void_p(stream) if stream else None, ctypes.c_void_p(nn),
                                            ctypes.c_void_p(lab), meta.ctypes.data_as(ctypes.c_void_p))
        if rc != ndbx.OK:
            raise ndbx.NdbxError(rc, ndbx.lib().ndbx_last_error().decode())
        return hip.download(nn, n, np.int32), hip.download(lab, n, np.int32), tuple(int(v) for v in meta)
    finally:
        hip.free(nn)
        hip.free(lab)


def with_context(ndbx, x_dev, n, d, eps, m, flags=0):
    with ndbx.Context(None, eps, m, flags=flags, device_ptr=x_dev, n=n, d=d) as c:
        c.classify().label()
        return c.nn(), c.labels(), c.meta()


def compare(ndbx, hip, x, eps, m, flags=0, what=""):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    xd = hip.upload(x)
    try:
        print(f"case noalloc {what} n={n} d={d}", flush=True)
        got = noalloc(ndbx, hip, xd, n, d, eps, m, flags)
        want = with_context(ndbx, xd, n, d, eps, m, flags)
    finally:
        hip.free(xd)
    np.testing.assert_array_equal(got[0], want[0], err_msg=what + " nn")
    np.testing.assert_array_equal(got[1], want[1], err_msg=what + " labels")
    assert got[2] == want[2], what
    return got


def test_noalloc_two_calls_second_smaller(gpu_lib, hip):
    x1, eps1, m1 = synth.make("blobs3d_10m", n=120_000)
    a = compare(gpu_lib, hip, x1, eps1, m1, what="first call, 3-D")
    assert a[2][0] > 0
    x2 = synth.blobs(30_000, 2, 12, 0.0, 60.0, None, 0.05, 909, sigma_range=(0.3, 1.5))
    b = compare(gpu_lib, hip, x2, synth.eps_f32("0.4"), 8, what="second call, smaller, 2-D")
    assert b[2][2] > 1
    # and the first input once more after the smaller one
    compare(gpu_lib, hip, x1, eps1, m1, what="third call, 3-D again")


@pytest.mark.parametrize("d", [1, 2, 3])
def test_noalloc_largest_size_sparse(gpu_lib, hip, d):
    """Every point in its own cell: the most cells, stencil windows and row
    slots the scratch array must hold."""
    n = gpu_lib.NOALLOC_MAX_POINTS
    rng = np.random.default_rng(50 + d)
    x = rng.uniform(0.0, 1000.0, size=(n, d)).astype(np.float32)
    eps = np.float32({1: 1e-3, 2: 0.05, 3: 0.5}[d])
    compare(gpu_lib, hip, x, eps, 2, what=f"largest sparse d={d}")


def test_noalloc_dense_2d_and_brute(gpu_lib, hip):
    x = synth.blobs(60_000, 2, 2, 0.0, 5.0, None, 0.0, 912, sigma_range=(0.05, 0.1))
    compare(gpu_lib, hip, x, np.float32(0.2), 50, what="dense 2-D")
    y = synth.blobs(12_000, 3, 6, 0.0, 40.0, None, 0.05, 913, sigma_range=(0.5, 2.0))
    compare(gpu_lib, hip, y, np.float32(1.0), 6, flags=gpu_lib.FORCE_BRUTE, what="all-pairs")


def test_noalloc_on_a_stream_and_empty(gpu_lib, hip):
    x, eps, m = synth.make("blobs2d_100k", n=50_000)
    xd = hip.upload(x)
    s = hip.stream()
    try:
        print("case noalloc on a stream", flush=True)
        got = noalloc(gpu_lib, hip, xd, len(x), 2, eps, m, stream=s)
        want = with_context(gpu_lib, xd, len(x), 2, eps, m)
    finally:
        hip.free(xd)
        hip.ok(hip.L.hipStreamDestroy(ctypes.c_void_p(s)))
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    assert got[2] == want[2]
    assert noalloc(gpu_lib, hip, 0, 0, 2, eps, m)[2] == (0, 0, 0)


def test_noalloc_refuses(gpu_lib, hip):
    n = gpu_lib.NOALLOC_MAX_POINTS + 1
    xd = hip.upload(np.zeros((n, 3), np.float32))
    try:
        with pytest.raises(gpu_lib.NdbxError) as e:
            noalloc(gpu_lib, hip, xd, n, 3, 0.5, 4)
        assert e.value.code == gpu_lib.EINVAL and "NDBX_NOALLOC_MAX_POINTS" in str(e.value)
        with pytest.raises(gpu_lib.NdbxError) as e:
            noalloc(gpu_lib, hip, xd, 100, 4, 0.5, 4)
        assert e.value.code == gpu_lib.EINVAL
        with pytest.raises(gpu_lib.NdbxError) as e:
            noalloc(gpu_lib, hip, xd, 100, 3, 0.5, 4, flags=gpu_lib.GRAPH)
        assert e.value.code == gpu_lib.EINVAL
    finally:
        hip.free(xd)
