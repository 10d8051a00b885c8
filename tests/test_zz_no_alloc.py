"""orbc_sort_pairs_noalloc: the sort for callers that cannot allocate device
memory, through the C ABI, against orbc_sort_pairs (the entry point whose
kernels it shares) on two calls in a row, the second smaller and with an odd
number of digit passes (the result comes back from the fixed scratch).

The device arrays are the caller's: this test allocates them with the HIP
runtime the library itself links (libamdhip64.so.7), as such a caller would.
"""
import ctypes as C

import numpy as np
import pytest

import openrbc_amd as orbc

pytestmark = pytest.mark.gpu

NOALLOC_SORT_MAX = 1 << 20


class _Dev:
    """Minimal HIP runtime access for caller-owned device buffers."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def put(self, dst, a):
        assert self.h.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # H2D

    def get(self, src, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), a.nbytes, 2) == 0   # D2H

    def free(self, p):
        self.h.hipFree(C.c_void_p(p))


def test_sort_pairs_noalloc_matches_sort_pairs():
    sim = orbc.Simulation(precision=orbc.MIXED)
    dev = _Dev()
    rng = np.random.default_rng(5)
    # 32 bits: four digit passes (result in the caller's arrays); 20 bits: three
    # passes (result copied back from the scratch); few distinct keys in the
    # second call, so stability is visible
    for n, bits, distinct in ((300000, 32, 1 << 32), (1000, 20, 37)):
        keys = (rng.integers(0, distinct, n, dtype=np.uint64) & ((1 << bits) - 1)).astype(np.uint32)
        vals = rng.permutation(n).astype(np.uint32)
        ek, ev = sim.sort_pairs(keys, vals, bits)
        dk, dv = dev.alloc(4 * n), dev.alloc(4 * n)
        try:
            dev.put(dk, keys)
            dev.put(dv, vals)
            sim.sort_pairs_device(dk, dv, n, bits)
            sim.synchronize()
            gk, gv = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
            dev.get(dk, gk)
            dev.get(dv, gv)
        finally:
            dev.free(dk)
            dev.free(dv)
        assert np.array_equal(gk, ek) and np.array_equal(gv, ev), (n, bits)
        assert np.all(np.diff(gk.astype(np.int64)) >= 0)


def test_sort_pairs_noalloc_rejects_sizes_beyond_the_scratch():
    sim = orbc.Simulation(precision=orbc.MIXED)
    dev = _Dev()
    p = dev.alloc(16)
    try:
        # rejected on the host before anything is enqueued
        with pytest.raises(orbc.OrbcError, match="ORBC_NOALLOC_SORT_MAX"):
            sim.sort_pairs_device(p, p, NOALLOC_SORT_MAX + 1, 32)
        with pytest.raises(orbc.OrbcError):
            sim.sort_pairs_device(p, p, 4, 33)
    finally:
        dev.free(p)
