"""Many searches in one launch: the slot engine of include/onion_gen_hip.h."""
import ctypes
from dataclasses import dataclass

from ._lib import OgError, err_buf, lib


class _Opts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("chains", ctypes.c_uint32),
                ("half", ctypes.c_uint32), ("rounds", ctypes.c_uint32),
                ("ring_cap", ctypes.c_uint32)]


class _Hit(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("offset", ctypes.c_uint64), ("slot", ctypes.c_uint32),
                ("packed", ctypes.c_uint8 * 32)]


@dataclass(frozen=True)
class BatchHit:
    tag: int
    slot: int
    offset: int
    packed: bytes


class Batch:
    def __init__(self, slots=64, chains=2048, half=256, rounds=1, ring_cap=0, device=0):
        h = ctypes.c_void_p()
        e = err_buf()
        rc = lib().og_batch_create(device, ctypes.byref(_Opts(slots, chains, half, rounds,
                                                               ring_cap)),
                                   ctypes.byref(h), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode())
        self._h = h
        self._out = (_Hit * (1 << 16))()

    def per_step(self):
        return lib().og_batch_per_step(self._h)

    def last_ms(self):
        return lib().og_batch_last_ms(self._h)

    def admit(self, start_xyzt, filters, tag):
        """Slot index of the admitted search, or None when every slot is busy."""
        slot = ctypes.c_int()
        e = err_buf()
        rc = lib().og_batch_admit(self._h, bytes(start_xyzt), filters.handle, tag,
                                  ctypes.byref(slot), e, len(e))
        if rc == -2:
            return None
        if rc != 0:
            raise OgError(e.value.decode())
        return slot.value

    def retire(self, slot):
        if lib().og_batch_retire(self._h, slot) != 0:
            raise OgError(f"could not retire slot {slot}")

    def step(self):
        n = ctypes.c_uint32()
        e = err_buf()
        rc = lib().og_batch_step(self._h, self._out, len(self._out), ctypes.byref(n), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode())
        return [BatchHit(self._out[i].tag, self._out[i].slot, self._out[i].offset,
                         bytes(self._out[i].packed)) for i in range(n.value)]

    def slot_state(self, slot):
        tag, examined, hits = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        active = ctypes.c_int()
        lib().og_batch_slot_state(self._h, slot, ctypes.byref(tag), ctypes.byref(examined),
                                  ctypes.byref(hits), ctypes.byref(active))
        return {"tag": tag.value, "examined": examined.value, "hits": hits.value,
                "active": bool(active.value)}

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


def footprint(filters):
    return lib().og_batch_footprint(filters.handle)
