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
    return buf.value.decode()


def get(key):
    """The entry under ``key`` as a dict, or None."""
    s = TuningShape()
    if lib().og_tuning_get(key.encode(), ctypes.byref(s)) != 1:
        return None
    return {"first": s.first, "second": s.second, "rate": s.rate, "at": s.at,
            "rounds": s.rounds, "radix": s.radix}


def put(key, first, second, rate, at=0, rounds=0, radix=0):
    return lib().og_tuning_put(key.encode(), ctypes.byref(
        TuningShape(first, second, rate, at, rounds, radix))) == 0
