"""The allocation-free re-expansion entry point (pbsam_xform_h_batch_noalloc) against
pbsam_xform_h_batch on the same pairs: two calls in a row on device buffers the test owns,
the second smaller (the scratch keeps the first call's contents), equal to the allocating
entry point bit for bit (the same kernels and launch shapes); sizes beyond the scratch and
orders outside [1, p] are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """device buffers from the HIP runtime the library itself is linked against"""

    def __init__(self):
        import pbsam_amd
        pbsam_amd.load()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def pairs(rng, n, p):
    from conftest import KAPPA  # noqa: F401
    d = rng.normal(size=(n, 3))
    d *= rng.uniform(9.0, 30.0, (n, 1)) / np.linalg.norm(d, axis=1)[:, None]
    a = rng.uniform(2.0, 6.0, (n, 2))
    order = rng.integers(1, p + 1, n).astype(np.int32)
    blocks = rng.normal(size=(n, p * p)) * 0.3 ** np.repeat(np.arange(p), 2 * np.arange(p) + 1)[None, :]
    return d, a, order, blocks


def noalloc(gpu, hip, d, a, order, blocks, p):
    L = gpu.L
    L.pbsam_xform_h_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    L.pbsam_xform_h_batch_noalloc.restype = ctypes.c_int
    n = len(order)
    bufs = [hip.put(np.ascontiguousarray(x)) for x in (d.reshape(-1), a.reshape(-1), order, blocks.reshape(-1))]
    out = hip.put(np.zeros(n * p * p))
    rc = L.pbsam_xform_h_batch_noalloc(gpu.h, n, p, *bufs, out)
    return rc, hip.get(out, np.zeros((n, p * p)))


@pytest.mark.parametrize("p", [10, 30])
def test_noalloc_equals_allocating_entry_point(gpu, p):
    hip = Hip()
    try:
        rng = np.random.default_rng(p)
        for n in (300, 37):   # the second call smaller than the first
            d, a, order, blocks = pairs(rng, n, p)
            rc, got = noalloc(gpu, hip, d, a, order, blocks, p)
            assert rc == 0, gpu.L.pbsam_last_error(gpu.h)
            ref = gpu.xform_h(d, a, order, blocks, p)
            assert np.array_equal(got, ref), (n, np.abs(got - ref).max())
            assert np.abs(ref).max() > 0
    finally:
        hip.free()


def test_noalloc_refuses_what_does_not_fit(gpu):
    import pbsam_amd
    hip = Hip()
    try:
        L = gpu.L
        L.pbsam_xform_h_batch_noalloc_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
        L.pbsam_xform_h_batch_noalloc_bytes.restype = ctypes.c_size_t
        p = 30
        n = 1
        while L.pbsam_xform_h_batch_noalloc_bytes(n + 1, p) <= (64 << 20):
            n += 1
        rng = np.random.default_rng(1)
        # one pair more than the scratch holds: refused before any launch (buffers of one pair
        # suffice, nothing is read)
        d, a, order, blocks = pairs(rng, 1, p)
        bufs = [hip.put(x) for x in (d.reshape(-1), a.reshape(-1), order, blocks.reshape(-1), np.zeros(p * p))]
        L.pbsam_xform_h_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
        assert L.pbsam_xform_h_batch_noalloc(gpu.h, n + 1, p, *bufs) == pbsam_amd.ERR_BADARG
        # an order beyond p: found on the device, reported as a bad argument
        order_bad = np.array([p + 3], dtype=np.int32)
        rc, _ = noalloc(gpu, hip, d, a, order_bad, blocks, p)
        assert rc == pbsam_amd.ERR_BADARG
        # and the scratch serves a good call afterwards
        rc, got = noalloc(gpu, hip, d, a, order, blocks, p)
        assert rc == 0 and np.array_equal(got, gpu.xform_h(d, a, order, blocks, p))
    finally:
        hip.free()
