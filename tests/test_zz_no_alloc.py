"""The allocation-free dissimilarity (kodama_hip_dissimilarity_inplace_noalloc)
against the allocating entry point kodama_hip_dissimilarity_inplace, through
the C ABI, on two calls in a row (the second smaller, so it reads scratch the
first left behind), plus the size limit of its static scratch. Device buffers
come from the HIP runtime the library itself links (ctypes), not from the
library."""
import ctypes as C

import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import _native as N

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libkodama_hip.so loaded."""

    def __init__(self):
        N.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _case(n, k, runs, seed):
    from kodama_amd import datasets as D
    x, _ = D.gaussian_mixture(n, 8, 5, seed=seed)
    g = K.KODAMA_knn_graph(x, k=k)
    gi, gd = g["indices"], g["distances"]
    rng = np.random.default_rng(seed)
    res = rng.integers(0, 5, size=(runs, n)).astype(np.int32)  # 0 = unassigned
    return gi, gd, res


def _noalloc(hip, gi, gd, res):
    n, k = gi.shape
    di, dd, dr = hip.put(gi), hip.put(gd), hip.put(res)
    N.call("kodama_hip_dissimilarity_inplace_noalloc", C.cast(di, C.POINTER(C.c_int32)),
           C.cast(dd, C.POINTER(C.c_float)), C.c_int32(n), C.c_int32(k), C.cast(dr, C.POINTER(C.c_int32)),
           C.c_int32(res.shape[0]), C.c_int32(0), C.c_void_p())
    hip.sync()
    return hip.get(di, gi), hip.get(dd, gd)


def test_noalloc_dissimilarity_equals_allocating_entry_point():
    hip = _Hip()
    try:
        for n, k, runs, seed in ((3000, 24, 37, 1), (700, 9, 5, 2)):
            gi, gd, res = _case(n, k, runs, seed)
            want_i, want_d = K.KODAMA_dissimilarity(gi, gd, res)
            got_i, got_d = _noalloc(hip, gi, gd, res)
            assert np.array_equal(got_i, want_i), (n, k, runs)
            assert np.array_equal(got_d, want_d), (n, k, runs)
    finally:
        hip.free()


def test_noalloc_dissimilarity_rejects_inputs_beyond_its_scratch():
    """samples x ceil4(runs) x 4 bytes above 64 MiB is EINVAL before any launch
    (no device memory is read: the pointers are never dereferenced)."""
    hip = _Hip()
    try:
        dummy = hip.put(np.zeros(4, np.int32))
        with pytest.raises(K.InvalidArgument):
            N.call("kodama_hip_dissimilarity_inplace_noalloc", C.cast(dummy, C.POINTER(C.c_int32)),
                   C.cast(dummy, C.POINTER(C.c_float)), C.c_int32(1 << 20), C.c_int32(4),
                   C.cast(dummy, C.POINTER(C.c_int32)), C.c_int32(17), C.c_int32(0), C.c_void_p())
    finally:
        hip.free()
