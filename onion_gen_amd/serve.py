"""The serve loop (og_serve_run): requests over time on one slot engine."""
import ctypes

from ._lib import OgError, err_buf, lib
from .batch import LEAST_SERVED, _Opts


class Request(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("seed", ctypes.c_uint8 * 32),
                ("filter", ctypes.c_char * 64), ("want_hits", ctypes.c_uint32)]


class _Found(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("packed", ctypes.c_uint8 * 32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("hits", ctypes.c_uint32),
                ("examined", ctypes.c_uint64), ("admitted_s", ctypes.c_double),
                ("first_hit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("found", _Found * 16), ("seen", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("keys_per_s", ctypes.c_double),
                ("device_busy_s", ctypes.c_double), ("examined", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("max_resident", ctypes.c_uint32),
                ("completed", ctypes.c_uint32),
                ("running", ctypes.c_uint32), ("queued", ctypes.c_uint32),
                ("rejected", ctypes.c_uint32), ("not_arrived", ctypes.c_uint32),
                ("p50_time_to_n", ctypes.c_double), ("p99_time_to_n", ctypes.c_double),
                ("p50_time_to_first", ctypes.c_double), ("p99_time_to_first", ctypes.c_double)]


STATUS = {0: "not_arrived", 1: "queued", 2: "running", 3: "done", 4: "rejected"}


def run(requests, duration, slots=64, chains=16384, half=256, rounds=1, max_queue=256,
        device_budget=0, policy=LEAST_SERVED, ring_cap=1 << 20, device=0):
    """``requests``: list of (arrival_s, seed32, filter_text, want_hits), sorted
    by arrival.  Returns (stats dict, list of result dicts)."""
    n = len(requests)
    reqs = (Request * max(n, 1))()
    for i, (t, seed, text, want) in enumerate(requests):
        reqs[i].arrival_s = t
        reqs[i].seed[:] = list(seed)
        reqs[i].filter = text.encode()
        reqs[i].want_hits = want
    results = (Result * max(n, 1))()
    stats = Stats()
    e = err_buf()
    opts = _Opts(slots, chains, half, rounds, ring_cap, policy)
    rc = lib().og_serve_run(device, ctypes.byref(opts), reqs, n, duration, max_queue,
                            device_budget, results, ctypes.byref(stats), e, len(e))
    if rc != 0:
        raise OgError(e.value.decode(errors="replace"))
    st = {f: getattr(stats, f) for f, _ in Stats._fields_}
    out = []
    for i in range(n):
        r = results[i]
        out.append({"status": STATUS[r.status], "hits": r.hits, "seen": r.seen,
                    "examined": r.examined,
                    "admitted_s": r.admitted_s, "first_hit_s": r.first_hit_s,
                    "done_s": r.done_s,
                    "found": [(r.found[j].offset, bytes(r.found[j].packed))
                              for j in range(min(r.hits, 16))]})
    return st, out
