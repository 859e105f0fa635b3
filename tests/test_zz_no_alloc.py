"""The allocation-free search (og_hip_search_noalloc) against the engine it is
a variant of, through the C ABI: for the same start point, literal prefix and
shape, its survivors over a window must be exactly those one advance of an
engine of that shape reports over the same window (bit for bit: offsets and
packed keys).  Two calls in a row, the second with a smaller shape from a
later candidate index, so the shared static workspace is checked for what the
first call left in it; then an overflow, which must be reported, and the
refusals.  Device memory for the outputs comes from the HIP runtime directly
(the entry point itself allocates nothing).
"""
import ctypes
import hashlib
import os

import pytest

from onion_gen_amd import filters, gpu, key
from onion_gen_amd._lib import err_buf, lib
from onion_gen_amd.gpu import _Survivor

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def hip():
    h = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                 "libamdhip64.so"))
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceOut:
    """`cap` og_survivor records and a count, in device memory."""

    def __init__(self, h, cap):
        self.h, self.cap = h, cap
        self.out, self.n = ctypes.c_void_p(), ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.out), max(cap, 1) * ctypes.sizeof(_Survivor)) == 0
        assert h.hipMalloc(ctypes.byref(self.n), 4) == 0

    def read(self):
        assert self.h.hipDeviceSynchronize() == 0
        n = ctypes.c_uint32()
        assert self.h.hipMemcpy(ctypes.byref(n), self.n, 4, 2) == 0  # device to host
        kept = min(n.value, self.cap, 16384)
        buf = (_Survivor * max(kept, 1))()
        if kept:
            assert self.h.hipMemcpy(buf, self.out, kept * ctypes.sizeof(_Survivor), 2) == 0
        return n.value, sorted((buf[i].offset // 8, bytes(buf[i].packed)) for i in range(kept))

    def free(self):
        self.h.hipFree(self.out)
        self.h.hipFree(self.n)


def noalloc(start, fs, k0, threads, half, dev):
    e = err_buf()
    rc = lib().og_hip_search_noalloc(0, bytes(start), k0, threads, half, fs.handle, dev.out,
                                     dev.cap, dev.n, None, e, len(e))
    return rc, e.value.decode()


def engine_window(start, fs, threads, half, k0):
    """The engine's survivors over [k0, k0 + batch): the advance that covers
    them (k0 a multiple of the batch)."""
    with gpu.Engine.with_filters(start, fs, threads=threads, half=half, rounds=1,
                                 hit_cap=1 << 16) as eng:
        batch = eng.batch()
        assert k0 % batch == 0
        for _ in range(k0 // batch):
            eng.advance()
        got = eng.advance()
    return sorted((s.offset // 8, s.packed) for s in got)


def test_noalloc_matches_the_engine_twice_then_overflows():
    h = hip()
    seed = hashlib.sha256(b"no alloc").digest()
    start = key.start_point(seed)
    fs = filters.FilterSet.parse("2x")
    assert lib().og_hip_search_noalloc_bytes(4096, 128) <= lib().og_hip_search_noalloc_max()
    dev = DeviceOut(h, 4096)
    try:
        # call 1: 4096 chains x 128 pairs (1,052,672 candidates) from k = 0
        say("case: noalloc 4096 x 128 from k = 0, prefix '2x'")
        rc, msg = noalloc(start, fs, 0, 4096, 128, dev)
        assert rc == 0, msg
        n1, got1 = dev.read()
        say("case: engine 4096 x 128, first advance")
        want1 = engine_window(start, fs, 4096, 128, 0)
        assert n1 == len(want1) and want1 and got1 == want1
        # call 2: a smaller shape, 1024 chains x 64 pairs, from its fourth batch
        k0 = 3 * 1024 * 129
        say(f"case: noalloc 1024 x 64 from k = {k0}")
        rc, msg = noalloc(start, fs, k0, 1024, 64, dev)
        assert rc == 0, msg
        n2, got2 = dev.read()
        say("case: engine 1024 x 64, fourth advance")
        want2 = engine_window(start, fs, 1024, 64, k0)
        assert n2 == len(want2) and want2 and got2 == want2
        assert all(k0 <= k < k0 + 1024 * 129 for k, _ in got2)
    finally:
        dev.free()
    # an output that holds fewer than were found: the count says so, and what
    # was written is the start of the same survivors
    small = DeviceOut(h, 8)
    try:
        say("case: noalloc 4096 x 128 from k = 0 into 8 records")
        rc, msg = noalloc(start, fs, 0, 4096, 128, small)
        assert rc == 0, msg
        n3, got3 = small.read()
        assert n3 == n1 > 8 and len(got3) == 8 and set(got3) <= set(want1)
    finally:
        small.free()


def test_noalloc_refusals():
    h = hip()
    start = key.start_point(hashlib.sha256(b"no alloc refusals").digest())
    dev = DeviceOut(h, 16)
    try:
        say("case: noalloc refusals (host-side checks only)")
        rc, msg = noalloc(start, filters.FilterSet.parse("suffix:aqd"), 0, 1024, 64, dev)
        assert rc != 0 and "one literal prefix" in msg
        rc, msg = noalloc(start, filters.FilterSet.parse(["2x", "3y"]), 0, 1024, 64, dev)
        assert rc != 0 and "one literal prefix" in msg
        rc, msg = noalloc(start, filters.FilterSet.parse("2x"), 0, 1 << 16, 1024, dev)
        assert rc != 0 and "workspace" in msg
        rc, msg = noalloc(start, filters.FilterSet.parse("2x"), 0, 1000, 64, dev)
        assert rc != 0 and "multiple of 256" in msg
    finally:
        dev.free()
