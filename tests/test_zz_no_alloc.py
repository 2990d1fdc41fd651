"""The allocation-free sort entry point (gs3d_radix_sort_noalloc) against the
existing one (gs3d_radix_sorter_sort), through the C ABI.

Device buffers for the caller's input and output come from the HIP runtime
the library itself loaded (hipMalloc through ctypes): the entry point
allocates nothing of its own. Two calls in a row, the second with fewer
pairs, must each equal the existing sorter's result bit for bit (keys and
values, so ties keep input order), leave the input unchanged and write
nothing past n; an in-place call (input = output) sorts the same way; more
pairs than the fixed scratch holds is an argument error raised before any
launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
SENTINEL = 0xDEADBEEF


def _case(msg):
    print(f"[case] {msg}", flush=True)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libgs3d.so uses."""

    def __init__(self, gs):
        gs.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "libgs3d.so did not load a HIP runtime"
        h = C.CDLL(path)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        self.h = h
        self.live = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.live.append(p.value)
        return p.value

    def upload(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.h.hipMemcpy(C.c_void_p(dst), C.c_void_p(a.ctypes.data), a.nbytes, H2D) == 0

    def download(self, src, n):
        a = np.empty(n, np.uint32)
        assert self.h.hipMemcpy(C.c_void_p(a.ctypes.data), C.c_void_p(src), a.nbytes, D2H) == 0
        return a

    def free_all(self):
        for p in self.live:
            self.h.hipFree(C.c_void_p(p))
        self.live = []


def _pairs(rng, n):
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    dup = rng.random(n) < 0.3  # a third of the keys from a small set: long runs of ties
    keys[dup] = rng.integers(0, 64, int(dup.sum()), dtype=np.uint32) * np.uint32(0x01010101)
    vals = (np.arange(n, dtype=np.uint32) * np.uint32(2654435761)) ^ np.uint32(0x5A5A5A5A)
    return keys, vals


def _existing(gs, device, keys, vals):
    s = gs.RadixSorter(device, len(keys))
    try:
        s.write(keys, vals)
        s.sort()
        return s.read()
    finally:
        s.close()


def test_noalloc_sort_equals_radix_sorter(gs, device):
    hip = _Hip(gs)
    rng = np.random.default_rng(11)
    cap = 3_000_017
    kin, vin, kout, vout = (hip.malloc(cap * 4 + 64) for _ in range(4))
    try:
        for call, n in enumerate((3_000_017, 250_001)):
            _case(f"noalloc sort call {call}: n={n}")
            keys, vals = _pairs(rng, n)
            hip.upload(kin, keys)
            hip.upload(vin, vals)
            hip.upload(kout, np.full(cap, SENTINEL, np.uint32))
            hip.upload(vout, np.full(cap, SENTINEL, np.uint32))
            gs.radix_sort_noalloc(device, kin, vin, kout, vout, n)
            device.poll()
            got_k, got_v = hip.download(kout, cap), hip.download(vout, cap)
            ek, ev = _existing(gs, device, keys, vals)
            assert np.array_equal(got_k[:n], ek), f"call {call}: keys differ from gs3d_radix_sorter_sort"
            assert np.array_equal(got_v[:n], ev), f"call {call}: values differ from gs3d_radix_sorter_sort"
            order = np.argsort(keys, kind="stable")
            assert np.array_equal(ev, vals[order]), f"call {call}: not a stable ascending sort"
            assert (got_k[n:] == SENTINEL).all() and (got_v[n:] == SENTINEL).all(), "wrote past n"
            assert np.array_equal(hip.download(kin, n), keys), "input keys changed"
            assert np.array_equal(hip.download(vin, n), vals), "input values changed"

        n = 65_537
        _case(f"noalloc sort in place: n={n}")
        keys, vals = _pairs(rng, n)
        hip.upload(kout, keys)
        hip.upload(vout, vals)
        gs.radix_sort_noalloc(device, kout, vout, kout, vout, n)
        device.poll()
        ek, ev = _existing(gs, device, keys, vals)
        assert np.array_equal(hip.download(kout, n), ek)
        assert np.array_equal(hip.download(vout, n), ev)

        _case("noalloc sort: zero pairs and too many pairs")
        gs.radix_sort_noalloc(device, kin, vin, kout, vout, 0)
        with pytest.raises(gs.Gs3dError, match="GS3D_RADIX_SORT_NOALLOC_MAX_KEYS"):
            gs.radix_sort_noalloc(device, kin, vin, kout, vout, gs.RADIX_SORT_NOALLOC_MAX_KEYS + 1)
        device.poll()
    finally:
        hip.free_all()
