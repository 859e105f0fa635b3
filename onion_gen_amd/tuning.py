"""The tuning cache (src/tuning.rs): measured shapes in the reference's
tuning.yaml, through og_tuning_path / og_tuning_get / og_tuning_put.

``first`` is pairs per window (half), ``second`` chains in flight, ``rate``
candidates per second, ``at`` unix seconds; ``rounds`` and ``radix`` extend
the reference's fields and are ignored by its reader."""
import ctypes

from ._lib import TuningShape, lib


def path():
    buf = ctypes.create_string_buffer(4096)
    if lib().og_tuning_path(buf, len(buf)) != 0:
        return None
    return buf.value.decode(errors="replace")


def get(key):
    """The entry under ``key`` as a dict, or None."""
    s = TuningShape()
    if lib().og_tuning_get(key.encode(), ctypes.byref(s)) != 1:
        return None
    return {"first": s.first, "second": s.second, "rate": s.rate, "at": s.at,
            "rounds": s.rounds, "radix": s.radix, "slots": s.slots}


def put(key, first, second, rate, at=0, rounds=0, radix=0, slots=0):
    return lib().og_tuning_put(key.encode(), ctypes.byref(
        TuningShape(first, second, rate, at, rounds, radix, slots))) == 0


def device_key(name, api="hip", threads=0):
    """src/tuning.rs:65-74 device_key, as og_tuning_device_key builds it."""
    n = lib().og_tuning_device_key(name.encode(), api.encode(), threads, None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().og_tuning_device_key(name.encode(), api.encode(), threads, buf, n + 1)
    return buf.value.decode(errors="replace")


def set_version(version):
    """The version the keys end in (the reference crate's 1.0.2 by default)."""
    return lib().og_tuning_set_version(version.encode()) == 0


def choose(rates, sizes, budget=0, tolerance=0.03):
    """og_tuning_choose over measured (rate, bytes) pairs: an index or None."""
    n = len(rates)
    r = (ctypes.c_double * max(n, 1))(*rates)
    b = (ctypes.c_uint64 * max(n, 1))(*sizes)
    i = lib().og_tuning_choose(r, b, n, budget, tolerance)
    return None if i >= n else i
