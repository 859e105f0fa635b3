"""Many searches in one launch: the slot engine of include/onion_gen_hip.h.

Resident searches keep their chain state and filter tables on the device; at
each step the policy (least served first, or round-robin) picks which of them
take the ``slots``."""
import ctypes
from dataclasses import dataclass

from ._lib import OgError, err_buf, lib


class _Opts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("chains", ctypes.c_uint32),
                ("half", ctypes.c_uint32), ("rounds", ctypes.c_uint32),
                ("ring_cap", ctypes.c_uint32), ("policy", ctypes.c_uint32)]


LEAST_SERVED, ROUND_ROBIN = 0, 1


class _Hit(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("offset", ctypes.c_uint64),
                ("search", ctypes.c_uint32), ("packed", ctypes.c_uint8 * 32)]


@dataclass(frozen=True)
class BatchHit:
    tag: int
    search: int
    offset: int
    packed: bytes


def tune(filter_sets, max_bytes=0, max_step_ms=0.0, device=0):
    """og_batch_tune: the unit shape a timed walk picks for this mix of filter
    sets (FilterSet objects) among those whose steps take at most
    `max_step_ms` (0: 25 ms), recorded in the tuning cache, as a dict with its
    aggregate rate."""
    n = len(filter_sets)
    arr = (ctypes.c_void_p * max(n, 1))(*[f.handle.value for f in filter_sets])
    best = _Opts()
    rate = ctypes.c_double()
    e = err_buf()
    if lib().og_batch_tune(device, arr, n, max_bytes, max_step_ms, ctypes.byref(best),
                           ctypes.byref(rate), e, len(e)) != 0:
        raise OgError(e.value.decode(errors="replace"))
    return {"slots": best.slots, "chains": best.chains, "half": best.half,
            "rate": rate.value,
            "step_ms": best.slots * best.chains * (2 * best.half + 1) / rate.value * 1e3
            if rate.value else None}


class Batch:
    def __init__(self, slots=64, chains=16384, half=256, rounds=1, ring_cap=0,
                 policy=LEAST_SERVED, device=0):
        h = ctypes.c_void_p()
        e = err_buf()
        rc = lib().og_batch_create(device, ctypes.byref(_Opts(slots, chains, half, rounds,
                                                               ring_cap, policy)),
                                   ctypes.byref(h), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        self._h = h
        self._out = (_Hit * (ring_cap or (1 << 16)))()

    def per_step(self):
        return lib().og_batch_per_step(self._h)

    def last_ms(self):
        return lib().og_batch_last_ms(self._h)

    def resident_bytes(self):
        return lib().og_batch_resident_bytes(self._h)

    def footprint(self, filters):
        return lib().og_batch_footprint(self._h, filters.handle)

    def admit(self, start_xyzt, filters, tag):
        """Makes a search resident; returns its id."""
        sid = ctypes.c_int()
        e = err_buf()
        rc = lib().og_batch_admit(self._h, bytes(start_xyzt), filters.handle, tag,
                                  ctypes.byref(sid), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        return sid.value

    def set_want(self, sid, want):
        """The search records at most `want` hits in all; later hits are
        counted (found) but not reported."""
        if lib().og_batch_set_want(self._h, sid, want) != 0:
            raise OgError(f"no resident search {sid}")

    def found(self, sid):
        return lib().og_batch_search_found(self._h, sid)

    def width(self, sid):
        """Unit slots the search takes per step (0 when not resident)."""
        return lib().og_batch_search_width(self._h, sid)

    def examined(self):
        """Candidates examined by every step so far, over all searches."""
        return lib().og_batch_examined(self._h)

    def retire(self, sid):
        if lib().og_batch_retire(self._h, sid) != 0:
            raise OgError(f"could not retire search {sid}")

    def step(self):
        n = ctypes.c_uint32()
        e = err_buf()
        rc = lib().og_batch_step(self._h, self._out, len(self._out), ctypes.byref(n), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        return [BatchHit(self._out[i].tag, self._out[i].search, self._out[i].offset,
                         bytes(self._out[i].packed)) for i in range(n.value)]

    def step_tags(self):
        """One step; the tag of every hit as a numpy array (no per-hit Python
        objects: for throughput runs whose searches hit often)."""
        import numpy as np
        n = ctypes.c_uint32()
        e = err_buf()
        rc = lib().og_batch_step(self._h, self._out, len(self._out), ctypes.byref(n), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        raw = np.frombuffer(self._out, dtype=np.uint64, count=n.value * (ctypes.sizeof(_Hit) // 8))
        return raw.reshape(n.value, ctypes.sizeof(_Hit) // 8)[:, 0].copy()

    def search_state(self, sid):
        tag, examined, hits = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        live = ctypes.c_int()
        lib().og_batch_search_state(self._h, sid, ctypes.byref(tag), ctypes.byref(examined),
                                    ctypes.byref(hits), ctypes.byref(live))
        return {"tag": tag.value, "examined": examined.value, "hits": hits.value,
                "live": bool(live.value)}

    def close(self):
        if self._h:
            lib().og_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
