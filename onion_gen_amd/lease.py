"""The lease adapter: the master/worker messages of src/proto.rs fed to a
device engine (og_hip_run_lease), so a GPU can serve as a fleet worker
without the Rust runtime.

``run(lease, filters, engine)`` takes a ``proto::Lease`` as its JSON object
({"lease", "seed" (hex), "first_block", "blocks", "expires_in"}) and returns
the ``proto::Found`` objects ({"lease", "block", "offset"}) and the
``proto::Heartbeat`` fields ({"lease", "blocks_done", "candidates"}) plus the
measured blocks per second a worker reports with its next ``AskWork``."""
import ctypes
import time

from ._lib import OgError, err_buf, lib


class _Lease(ctypes.Structure):
    _fields_ = [("lease", ctypes.c_uint64), ("seed", ctypes.c_uint8 * 32),
                ("first_block", ctypes.c_uint64), ("blocks", ctypes.c_uint64)]


class _Find(ctypes.Structure):
    _fields_ = [("lease", ctypes.c_uint64), ("block", ctypes.c_uint64),
                ("offset", ctypes.c_uint64)]


def parse_seed(text):
    """The root seed of a lease: 64 hex digits (src/worker/run.rs parse_seed)."""
    raw = bytes.fromhex(text.strip())
    if len(raw) != 32:
        raise ValueError(f"a seed is 32 bytes, this one is {len(raw)}")
    return raw


def run(lease, filters, engine, launches=0, cap=1 << 20):
    """Searches the lease's blocks with ``engine`` (a gpu.Engine built with
    ``filters``): returns {"found": [Found...], "heartbeat": Heartbeat,
    "blocks_per_second": float}."""
    lz = _Lease(lease["lease"], (ctypes.c_uint8 * 32)(*parse_seed(lease["seed"])),
                lease["first_block"], lease["blocks"])
    out = (_Find * cap)()
    n, cand, done = ctypes.c_size_t(), ctypes.c_uint64(), ctypes.c_uint64()
    e = err_buf()
    t0 = time.perf_counter()
    rc = lib().og_hip_run_lease(engine._h, ctypes.byref(lz), filters.handle, launches, out, cap,
                                ctypes.byref(n), ctypes.byref(cand), ctypes.byref(done), e, len(e))
    elapsed = time.perf_counter() - t0
    if rc != 0:
        raise OgError(e.value.decode(errors="replace"))
    found = [{"lease": out[i].lease, "block": out[i].block, "offset": out[i].offset}
             for i in range(n.value)]
    return {"found": found,
            "heartbeat": {"lease": lease["lease"], "blocks_done": done.value,
                          "candidates": cand.value},
            "blocks_per_second": done.value / elapsed if elapsed > 0 else None}
