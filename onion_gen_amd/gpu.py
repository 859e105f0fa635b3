"""The device engine behind the reference's gpu::Engine boundary.

Mirrors src/gpu/mod.rs: ``Survivor`` (:113-119), ``hit_room`` (:306-312),
``Engine.new`` (:346-353), ``threads``/``batch`` (:446-467) and ``advance``
(:470-476), over the C ABI of include/onion_gen_hip.h.
"""
import ctypes
from dataclasses import dataclass

from ._lib import OgError, err_buf, lib


class _Survivor(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("packed", ctypes.c_uint8 * 32)]


class _Opts(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("half", ctypes.c_uint32),
                ("rounds", ctypes.c_uint32), ("hit_cap", ctypes.c_uint32),
                ("sign_stage", ctypes.c_uint32), ("max_bytes", ctypes.c_uint64)]


@dataclass(frozen=True)
class Survivor:
    """A candidate the device could not rule out (src/gpu/mod.rs:113-119).

    ``offset`` is the scalar delta from the start scalar (8k for candidate k);
    ``packed`` is the canonical y, sign bit clear.
    """
    offset: int
    packed: bytes


@dataclass(frozen=True)
class Device:
    index: int
    name: str
    memory: int
    compute_units: int


def device_count():
    return lib().og_hip_device_count()


def look():
    """Devices the HIP path reaches (src/gpu/mod.rs:165-244 for this API)."""
    out = []
    for i in range(device_count()):
        name = ctypes.create_string_buffer(256)
        mem = ctypes.c_uint64()
        cus = ctypes.c_int()
        if lib().og_hip_device_info(i, name, 256, ctypes.byref(mem), ctypes.byref(cus)) == 0:
            out.append(Device(i, name.value.decode(errors="replace"), mem.value, cus.value))
    return out


def hit_room(batch, bitmap_bits):
    return lib().og_hit_room(batch, bitmap_bits)


class Engine:
    """One device, its chains and its buffers.

    ``Engine.new(start, threads, bitmap, bitmap_bits)`` is the reference's
    constructor; ``Engine.with_filters(start, filters, ...)`` is the extended one
    that takes a compiled :class:`onion_gen_amd.filters.FilterSet`.
    """

    def __init__(self, handle):
        self._h = handle
        self._out = None
        self._room = lib().og_hip_engine_room(handle)

    @classmethod
    def new(cls, start_xyzt, threads, bitmap, bitmap_bits, device=0):
        words = (ctypes.c_uint64 * len(bitmap))(*bitmap)
        h = ctypes.c_void_p()
        e = err_buf()
        rc = lib().og_hip_engine_create(device, bytes(start_xyzt), threads, words, len(bitmap),
                                        bitmap_bits, ctypes.byref(h), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        return cls(h)

    @classmethod
    def with_filters(cls, start_xyzt, filters, threads=0, half=0, rounds=0, hit_cap=0, device=0,
                     sign_stage=0, max_bytes=0):
        """``sign_stage`` (og_engine_opts): 0 chosen from the filters, 1 the
        search loop, 2 the address kernel.  ``max_bytes``: with threads = 0,
        the device memory the chosen shape may hold (0: half the free memory)."""
        opts = _Opts(threads, half, rounds, hit_cap, sign_stage, max_bytes)
        h = ctypes.c_void_p()
        e = err_buf()
        rc = lib().og_hip_engine_create_ex(device, bytes(start_xyzt), ctypes.byref(opts),
                                           filters.handle, ctypes.byref(h), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        return cls(h)

    @staticmethod
    def tune(start_xyzt, filters, device=0, max_bytes=0):
        """Measures the walk of shapes within ``max_bytes`` of device memory
        (og_hip_engine_tune; 0: half the free memory), records the shape it
        keeps in the tuning cache and returns it as a dict."""
        from ._lib import TuningShape
        best = TuningShape()
        e = err_buf()
        if lib().og_hip_engine_tune(device, bytes(start_xyzt), filters.handle, max_bytes,
                                    ctypes.byref(best), e, len(e)) != 0:
            raise OgError(e.value.decode(errors="replace"))
        return {"threads": best.second, "half": best.first, "rounds": best.rounds,
                "rate": best.rate, "at": best.at}

    def bytes(self):
        """Device memory this engine holds (og_hip_engine_bytes)."""
        return lib().og_hip_engine_bytes(self._h)

    def shape(self):
        """(threads, half, rounds) this engine runs."""
        t, h, r = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        lib().og_hip_engine_shape(self._h, ctypes.byref(t), ctypes.byref(h), ctypes.byref(r))
        return t.value, h.value, r.value

    def sign_stage(self):
        """Where this engine forms the sign of x: 1 the search loop, 2 the
        address kernel (og_hip_engine_sign_stage)."""
        return lib().og_hip_engine_sign_stage(self._h)

    def kernels(self):
        """The search kernels this engine runs: 1 the key forms, 2 the pattern
        forms (og_hip_engine_kernels)."""
        return lib().og_hip_engine_kernels(self._h)

    def restart(self, start_xyzt):
        """Starts over from another start point (og_hip_engine_restart)."""
        e = err_buf()
        if lib().og_hip_engine_restart(self._h, bytes(start_xyzt), e, len(e)) != 0:
            raise OgError(e.value.decode(errors="replace"))

    def room(self):
        """Survivors one advance may return (og_hip_engine_room)."""
        return self._room

    def threads(self):
        return lib().og_hip_engine_threads(self._h)

    def batch(self):
        return lib().og_hip_engine_batch(self._h)

    def last_ms(self):
        return lib().og_hip_engine_last_ms(self._h)

    def advance(self, cap=0):
        """Examines the next ``batch()`` candidates; returns the survivors
        (room for ``cap`` of them; by default the engine's survivor room)."""
        cap = cap or self._room
        if self._out is None or len(self._out) < cap:
            self._out = (_Survivor * cap)()
        n = ctypes.c_uint32()
        e = err_buf()
        rc = lib().og_hip_engine_advance(self._h, self._out, cap, ctypes.byref(n), e, len(e))
        if rc != 0:
            raise OgError(e.value.decode(errors="replace"))
        return [Survivor(self._out[i].offset, bytes(self._out[i].packed)) for i in range(n.value)]

    def close(self):
        if self._h:
            lib().og_hip_engine_destroy(self._h)
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
