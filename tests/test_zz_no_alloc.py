"""The allocation-free record sort (gpq_sort_events_noalloc) against the allocating one
(gpq_selftest_sort) on the same records, through the C ABI: two calls in a row, the second
with fewer records, then in place, then a size above the scratch limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EVENT = np.dtype([("ts", "<u4"), ("dst", "<u4"), ("src", "<u4"), ("seq", "<u4")])


def _events(rng, n):
    ev = np.zeros(n, dtype=EVENT)
    ev["ts"] = np.float32(rng.exponential(3.0, n)).view(np.uint32)
    ev["ts"][: n // 4] = ev["ts"][n // 2: n // 2 + n // 4]   # timestamp ties: the order falls to (src, seq)
    ev["dst"] = rng.integers(0, 1 << 20, n)
    ev["src"] = rng.integers(0, 64, n)
    ev["seq"] = rng.integers(0, 1 << 31, n)
    return ev


def _selftest_sort(gpq, ev):
    out = np.empty_like(ev)
    gpq._lib.check(gpq.lib.gpq_selftest_sort(ev.ctypes.data_as(ctypes.c_void_p), len(ev),
                                             out.ctypes.data_as(ctypes.c_void_p)), "selftest_sort")
    return out


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself runs on."""

    def __init__(self):
        import os
        path = "libamdhip64.so"
        for cand in ("/opt/rocm/lib/libamdhip64.so",):
            if os.path.exists(cand):
                path = cand
        self.lib = ctypes.CDLL(path)
        for fn in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.lib, fn).restype = ctypes.c_int
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def free(self, p):
        assert self.lib.hipFree(p) == 0

    def to_device(self, p, arr):
        assert self.lib.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0    # host to device

    def to_host(self, arr, p):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), p, arr.nbytes, 2) == 0    # device to host

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


def _noalloc(gpq, hip, ev, in_place=False):
    """gpq_sort_events_noalloc on device copies of ev; returns (sorted records, input after)."""
    n = len(ev)
    d_in = hip.alloc(ev.nbytes)
    d_out = d_in if in_place else hip.alloc(ev.nbytes)
    try:
        hip.to_device(d_in, ev)
        if not in_place:
            hip.to_device(d_out, np.full_like(ev, 0xFF))
        hip.sync()
        gpq._lib.check(gpq.lib.gpq_sort_events_noalloc(d_in, n, d_out, None), "sort_events_noalloc")
        hip.sync()
        got = np.empty_like(ev)
        after = np.empty_like(ev)
        hip.to_host(got, d_out)
        hip.to_host(after, d_in)
        return got, after
    finally:
        hip.free(d_in)
        if not in_place:
            hip.free(d_out)


def test_noalloc_sort_equals_allocating_sort(gpq):
    hip = _Hip()
    rng = np.random.default_rng(7)
    # 300000 records take 8 merge passes (the run ends in the first buffer), 5000 take 2, 3000
    # take 1 (it ends in the partner's place), 700 none
    for n in (300000, 5000, 3000, 700):
        ev = _events(rng, n)
        ref = _selftest_sort(gpq, ev)
        got, after = _noalloc(gpq, hip, ev)
        assert np.array_equal(got, ref), n
        assert np.array_equal(after, ev)                                           # input untouched
    keys = np.lexsort((ref["seq"], ref["src"], ref["ts"]))
    assert np.array_equal(ref, ref[keys])


def test_noalloc_sort_in_place_and_limits(gpq):
    hip = _Hip()
    rng = np.random.default_rng(11)
    for n in (1 << 20, 4097):                  # the largest size the scratch array holds, then a smaller one
        ev = _events(rng, n)
        ref = _selftest_sort(gpq, ev)
        got, _ = _noalloc(gpq, hip, ev, in_place=True)
        assert np.array_equal(got, ref), n
    small = hip.alloc(256)
    try:
        st = gpq.lib.gpq_sort_events_noalloc(small, (1 << 20) + 1, small, None)
        assert st == 2                                   # GPQ_ERR_CAPACITY, nothing launched
        assert gpq.lib.gpq_sort_events_noalloc(small, 0, small, None) == 0
        hip.sync()
    finally:
        hip.free(small)
