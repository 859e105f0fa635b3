"""Compiled filter sets (src/filter.rs FilterSet) over the C ABI."""
import ctypes

from ._lib import OgError, err_buf, lib


class FilterSet:
    """``FilterSet.parse(lines, index_bits=0)``: parse, dedupe, absorb, index.

    Mirrors FilterSet::parse_all / new / with_index_bits (src/filter.rs:1930,
    1326, 1344), ``index_words`` (:1857), ``match_key`` (:2098) and
    ``match_whole_address`` (:1673).
    """

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def parse(cls, lines, index_bits=0):
        if isinstance(lines, str):
            lines = [lines]
        text = "\n".join(lines).encode()
        e = err_buf()
        h = lib().og_filters_parse(text, index_bits, e, len(e))
        if not h:
            raise OgError(e.value.decode(errors="replace"))
        return cls(ctypes.c_void_p(h))

    def __len__(self):
        return lib().og_filters_count(self.handle)

    def index_words(self):
        """(list of u64 words, bits) or None when no index was built."""
        bits = lib().og_filters_index_words(self.handle, None, 0)
        if not bits:
            return None
        n = ((1 << bits) + 63) // 64
        buf = (ctypes.c_uint64 * n)()
        lib().og_filters_index_words(self.handle, buf, n)
        return list(buf), bits

    def match_key(self, packed):
        return lib().og_filters_match_key(self.handle, bytes(packed))

    def match_address(self, address_bytes):
        return lib().og_filters_match_address(self.handle, bytes(address_bytes))

    def describe(self):
        """The structures the device reads, one line each (FilterSet::
        describe_structures, src/filter.rs:1700)."""
        n = lib().og_filters_describe(self.handle, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().og_filters_describe(self.handle, buf, n + 1)
        return buf.value.decode(errors="replace").splitlines()

    def scan(self, packed, sign):
        """The device matcher of the general forms and the expressions, run on
        the host for one key: (key_hit, need, hit) as og_filters_scan_host
        defines them."""
        kh, nd, hit = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        if lib().og_filters_scan_host(self.handle, bytes(packed), sign, ctypes.byref(kh),
                                      ctypes.byref(nd), ctypes.byref(hit)) != 0:
            raise OgError("the set has neither general filter nor expression, or does not compile")
        return bool(kh.value), bool(nd.value), bool(hit.value)

    def scan_pair(self, a, b):
        """The key stage for two keys at once, as the kernels run the two
        candidates of a pair: ((key_hit_a, need_a), (key_hit_b, need_b))."""
        kh, nd = (ctypes.c_int * 2)(), (ctypes.c_int * 2)()
        if lib().og_filters_scan_pair_host(self.handle, bytes(a), bytes(b), kh, nd) != 0:
            raise OgError("the set runs no automaton bank or expression")
        return (bool(kh[0]), bool(nd[0])), (bool(kh[1]), bool(nd[1]))

    def exact_table(self, packed):
        """The device's exact table behind the bitmap, run on the host for one
        canonical key: whether a prefix of some length in the set matches its
        first 64 bits."""
        hit = ctypes.c_int()
        if lib().og_filters_exact_table_host(self.handle, bytes(packed), ctypes.byref(hit)) != 0:
            raise OgError("the set has no exact table (not a dictionary of prefixes)")
        return bool(hit.value)

    def address_window(self, address):
        """The address stage on a 35-byte address: (window, match) - whether
        the address-stage window prefilter lets the automata walk (always,
        without one), and whether they accept the address."""
        window, match = ctypes.c_int(), ctypes.c_int()
        if lib().og_filters_address_window_host(self.handle, bytes(address), ctypes.byref(window),
                                                ctypes.byref(match)) != 0:
            raise OgError("the set runs no automaton")
        return bool(window.value), bool(match.value)

    def close(self):
        if self.handle:
            lib().og_filters_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
