"""The allocation-free event sort (nh_sort_events_noalloc) against the
allocating one (nh_debug_sort_events) and numpy's lexsort.

Both sort a round's speciation events by the PackedEvent order
(necsim/core/src/event.rs:256-280).  The allocation-free variant reads and
writes caller-owned device buffers and keeps its sort buffers in the
library's file-scope scratch array, so two calls in a row (the second one
smaller) must each give the allocating sort's bytes: nothing of the first
call's scratch contents may leak into the second.  Device buffers for the
test come from the HIP runtime directly (hipMalloc / hipMemcpy through
ctypes).
"""
import ctypes

import numpy as np
import pytest

from necsim_hip import EVENT_DTYPE, _ffi

pytestmark = pytest.mark.gpu

MAX_EVENTS = 1 << 20
H2D, D2H = 1, 2


def announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 1)) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(arr.ctypes.data, self.p, arr.nbytes, D2H) == 0

    def free(self):
        self.hip.hipFree(self.p)


def events_with_ties(rng, n):
    """speciation events with runs of equal times: coalesced duplicates that
    differ only in reference and prior time, and runs that mix origins"""
    ev = np.zeros(n, EVENT_DTYPE)
    ev["global_reference"] = rng.permutation(n).astype(np.uint64) + 1
    ev["event_time"] = rng.integers(1, max(n // 3, 2), n).astype(np.float64) * 0.5
    ev["prior_time"] = rng.integers(0, 8, n).astype(np.float64) * 0.25
    ev["x"] = rng.integers(0, 4, n)
    ev["y"] = rng.integers(0, 4, n)
    ev["index"] = rng.integers(0, 2, n)
    return ev


def expected(ev):
    order = np.lexsort((ev["global_reference"], ev["prior_time"], ev["index"], ev["y"], ev["x"],
                        ev["event_time"]))
    return ev[order]


def allocating_sort(ctx, ev, device_sized):
    out = np.zeros_like(ev)
    _ffi.check(_ffi.lib().nh_debug_sort_events(ctx.handle, _ffi.ptr(ev), len(ev), int(device_sized),
                                                _ffi.ptr(out)))
    return out


def test_noalloc_sort_matches_allocating_sort_twice(ctx, hip):
    rng = np.random.default_rng(23)
    d_in = DeviceBuffer(hip, MAX_EVENTS * EVENT_DTYPE.itemsize)
    d_out = DeviceBuffer(hip, MAX_EVENTS * EVENT_DTYPE.itemsize)
    try:
        for device_sized in (0, 1):
            for n in (300_000, 1_000):  # the second call is smaller than the first
                announce(f"noalloc sort: n={n} device_sized={device_sized}")
                ev = events_with_ties(rng, n)
                ref = allocating_sort(ctx, ev, device_sized)
                d_in.upload(ev)
                _ffi.check(_ffi.lib().nh_sort_events_noalloc(ctx.handle, d_in.p, n, device_sized, d_out.p))
                got = np.zeros_like(ev)
                d_out.download(got)
                assert got.tobytes() == ref.tobytes(), (n, device_sized)
                assert got.tobytes() == expected(ev).tobytes(), (n, device_sized)
    finally:
        d_in.free()
        d_out.free()


def test_noalloc_sort_limits(ctx, hip):
    announce("noalloc sort: size limit and argument checks (no kernel runs)")
    d = DeviceBuffer(hip, 64)
    try:
        rc = _ffi.lib().nh_sort_events_noalloc(ctx.handle, d.p, MAX_EVENTS + 1, 0, ctypes.c_void_p(d.p.value + 8))
        assert rc == -7  # NH_E_CAPACITY
        assert "at most" in _ffi.last_error()
        assert _ffi.lib().nh_sort_events_noalloc(ctx.handle, d.p, 1, 0, d.p) == -1  # in == out
        assert _ffi.lib().nh_sort_events_noalloc(ctx.handle, None, 0, 0, None) == 0  # empty
    finally:
        d.free()
