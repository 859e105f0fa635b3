"""ctypes binding of the CPU oracle (oracle/liboracle.so): the checker the
parity tests compare the HIP path against.  Builds it with make when absent."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")


class OHit(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint64), ("pk", ctypes.c_uint8 * 32),
                ("address", ctypes.c_char * 57), ("filter", ctypes.c_int32)]


_lib = None


def load(path=None):
    """The oracle library; `path` loads another build of the same sources (the
    natively built timing baseline) with the same declarations."""
    global _lib
    if path is not None:
        return _declare(ctypes.CDLL(path))
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    _lib = _declare(ctypes.CDLL(LIB))
    return _lib


def _declare(L):
    c = ctypes
    L.ofilter_parse.restype = c.c_void_p
    L.ofilter_parse.argtypes = [c.c_char_p, c.c_char_p, c.c_size_t]
    L.ofilter_free.argtypes = [c.c_void_p]
    L.ofilter_count.restype = c.c_size_t
    L.ofilter_count.argtypes = [c.c_void_p]
    L.ofilter_canonical.restype = c.c_char_p
    L.ofilter_canonical.argtypes = [c.c_void_p, c.c_size_t]
    L.ofilter_index.restype = c.c_uint32
    L.ofilter_index.argtypes = [c.c_void_p, c.c_uint32, c.c_void_p, c.c_size_t]
    L.ofilter_match_address.argtypes = [c.c_void_p, c.c_char_p]
    L.ofilter_match_key.argtypes = [c.c_void_p, c.c_char_p]
    L.osearch.restype = c.c_size_t
    L.osearch.argtypes = [c.c_char_p, c.c_uint64, c.c_uint64, c.c_size_t, c.c_void_p,
                          c.c_void_p, c.c_size_t]
    L.osearch_mt.restype = c.c_size_t
    L.osearch_mt.argtypes = [c.c_char_p, c.c_uint64, c.c_uint64, c.c_size_t, c.c_void_p,
                             c.c_int, c.c_void_p, c.c_size_t]
    L.odigest.argtypes = [c.c_char_p, c.c_uint64, c.c_uint64, c.c_size_t, c.c_void_p]
    L.obench_batch.restype = c.c_double
    L.obench_batch.argtypes = [c.c_char_p, c.c_size_t, c.c_int, c.c_double, c.c_void_p,
                               c.POINTER(c.c_uint64)]
    L.ochacha_stream.argtypes = [c.c_uint64, c.c_char_p, c.c_size_t]
    L.orun_block_seed.argtypes = [c.c_char_p, c.c_uint64, c.c_char_p]
    L.okey_add_to_scalar.argtypes = [c.c_char_p, c.c_uint64]
    L.okey_expanded_secret_at_offset.argtypes = [c.c_char_p, c.c_uint64, c.c_char_p]
    L.opoint_from_bytes.argtypes = [c.c_void_p, c.c_char_p]
    L.obatch_new.restype = c.c_void_p
    L.obatch_new.argtypes = [c.c_size_t]
    L.obatch_free.argtypes = [c.c_void_p]
    L.obatch_run.argtypes = [c.c_void_p, c.c_void_p]
    L.obatch_packed.restype = c.POINTER(c.c_uint8)
    L.obatch_packed.argtypes = [c.c_void_p, c.c_size_t]
    L.obatch_packed_with_sign.argtypes = [c.c_void_p, c.c_size_t, c.c_char_p]
    L.obench_pairs.restype = c.c_double
    L.obench_pairs.argtypes = [c.c_char_p, c.c_size_t, c.c_int, c.c_double, c.c_void_p,
                               c.POINTER(c.c_uint64)]
    L.opairs_new.restype = c.c_void_p
    L.opairs_new.argtypes = [c.c_size_t]
    L.opairs_free.argtypes = [c.c_void_p]
    L.opairs_batch_size.restype = c.c_size_t
    L.opairs_batch_size.argtypes = [c.c_void_p]
    L.opairs_begin.argtypes = [c.c_void_p, c.c_void_p]
    L.opairs_run.restype = c.c_size_t
    L.opairs_run.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p]
    L.opairs_survivor.argtypes = [c.c_void_p, c.c_size_t, c.POINTER(c.c_uint32), c.c_char_p]
    L.opairs_delta.restype = c.c_uint64
    L.opairs_delta.argtypes = [c.c_void_p, c.c_uint32]
    for name in ("oracle_fe_mul", "oracle_fe_add", "oracle_fe_sub"):
        getattr(L, name).argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.c_size_t]
    for name in ("oracle_fe_invert", "oracle_fe_canon"):
        getattr(L, name).argtypes = [c.c_char_p, c.c_char_p, c.c_size_t]
    return L


class OracleFilters:
    def __init__(self, lines):
        L = load()
        text = "\n".join([lines] if isinstance(lines, str) else lines).encode()
        err = ctypes.create_string_buffer(256)
        self.h = L.ofilter_parse(text, err, 256)
        if not self.h:
            raise ValueError(err.value.decode())

    def __len__(self):
        return load().ofilter_count(self.h)

    def canonical(self, i):
        return load().ofilter_canonical(self.h, i).decode()

    def index(self, bits=0):
        L = load()
        b = L.ofilter_index(self.h, bits, None, 0)
        if not b:
            return None
        n = ((1 << b) + 63) // 64
        buf = (ctypes.c_uint64 * n)()
        L.ofilter_index(self.h, bits, buf, n)
        return list(buf), b

    def match_address(self, ab):
        return load().ofilter_match_address(self.h, bytes(ab))

    def match_key(self, packed):
        return load().ofilter_match_key(self.h, bytes(packed))

    def __del__(self):
        try:
            load().ofilter_free(self.h)
        except Exception:
            pass


def start_point(seed):
    out = ctypes.create_string_buffer(128)
    load().okey_start_point(bytes(seed), out)
    return out.raw


def public_key(seed):
    out = ctypes.create_string_buffer(32)
    load().okey_public_key(bytes(seed), out)
    return out.raw


def address(pk):
    out = ctypes.create_string_buffer(57)
    load().okey_address(bytes(pk), out)
    return out.value.decode()


def search(start, k0, n, width, filters, cap=1 << 16):
    """Oracle hits for k0 <= k < k0+n: sorted (k, pk, address)."""
    buf = (OHit * cap)()
    cnt = load().osearch(bytes(start), k0, n, width, filters.h, buf, cap)
    if cnt > cap:
        raise ValueError(f"{cnt} oracle hits exceed the cap {cap}")
    return sorted((buf[i].k, bytes(buf[i].pk), buf[i].address.decode()) for i in range(cnt))


def search_mt(start, k0, n, width, filters, threads=None, cap=1 << 16):
    """search() split over threads (one contiguous part each): the same hits."""
    threads = threads or min(16, os.cpu_count() or 1)
    buf = (OHit * cap)()
    cnt = load().osearch_mt(bytes(start), k0, n, width, filters.h, threads, buf, cap)
    if cnt > cap:
        raise ValueError(f"{cnt} oracle hits exceed the cap {cap}")
    return sorted((buf[i].k, bytes(buf[i].pk), buf[i].address.decode()) for i in range(cnt))


def digest(start, k0, n, width=256):
    out = (ctypes.c_uint64 * 2)()
    load().odigest(bytes(start), k0, n, width, out)
    return out[0], out[1]


def batch_chain(start, width, batches):
    """Packed keys with sign from a BatchEngine of ``width``, ``batches`` runs."""
    L = load()
    e = L.obatch_new(width)
    acc = ctypes.create_string_buffer(bytes(start), 128)
    # obatch_run takes an opoint (4 x 4 x u64 limbs); build it from bytes
    pt = (ctypes.c_uint64 * 16)()
    L.opoint_from_bytes(pt, acc)
    out = []
    for _ in range(batches):
        L.obatch_run(e, pt)
        for i in range(width):
            b = ctypes.create_string_buffer(32)
            L.obatch_packed_with_sign(e, i, b)
            out.append(b.raw)
    L.obatch_free(e)
    return out


def pair_engine_hits(start, batch, rounds, filters):
    """The PairEngine restatement over `rounds` rounds from `start` (the point
    of delta 0): (sorted (k, packed y) of the survivors that match the set's
    key filters, the set of every delta the rounds covered)."""
    L = load()
    acc = ctypes.create_string_buffer(4 * 4 * 8)  # opoint: 4 field elements of 4 u64
    L.opoint_from_bytes(acc, bytes(start))
    e = L.opairs_new(batch)
    L.opairs_begin(e, acc)
    n = L.opairs_batch_size(e)
    hits, deltas = [], set()
    idx = ctypes.c_uint32()
    packed = ctypes.create_string_buffer(32)
    for _ in range(rounds):
        k = L.opairs_run(e, acc, filters.h)
        for i in range(n):
            deltas.add(L.opairs_delta(e, i))
        for i in range(k):
            L.opairs_survivor(e, i, ctypes.byref(idx), packed)
            if L.ofilter_match_key(filters.h, packed.raw) >= 0:
                hits.append((L.opairs_delta(e, idx.value) // 8, packed.raw))
    L.opairs_free(e)
    return sorted(hits), deltas
