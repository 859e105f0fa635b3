"""The filter rules the device path shares with the reference, pinned against
independent statements of them (not against the oracle's parser, which follows
the same reading of src/filter.rs).

- The prefix index (src/filter.rs:2248-2299, 2301-2350, 2321-2604): among the
  widths 16..24 the narrowest whose measured occupancy is at most 1%, else the
  widest; discarded above 25% occupancy.  A prefix of m symbols fills the slots
  of its 5m leading bits: one slot when 5m >= bits, a run of 2^(bits - 5m)
  otherwise.  Slot s is bit s & 63 of word s >> 6.  Restated here in Python
  over exact slot sets, and compared with og_filters_index_words bit for bit.
- Absorption and matching, with the reference's own test cases
  (src/filter.rs:3468-3486).
"""
import random

from onion_gen_amd import filters, key, workload

ALPHABET = workload.ALPHABET


def prefix_value(text):
    v = 0
    for c in text:
        v = (v << 5) | ALPHABET.index(c)
    return v, 5 * len(text)


def slots(prefixes, bits):
    out = set()
    for p in prefixes:
        v, n = prefix_value(p)
        if n >= bits:
            out.add(v >> (n - bits))
        else:
            lo = v << (bits - n)
            out.update(range(lo, lo + (1 << (bits - n))))
    return out


def reference_index(prefixes):
    """(bits, slot set) of src/filter.rs's automatic choice, or None."""
    chosen = None
    for bits in range(16, 25):
        s = slots(prefixes, bits)
        chosen = (bits, s)
        if len(s) / (1 << bits) <= 0.01:
            break
    bits, s = chosen
    if len(s) / (1 << bits) > 0.25:
        return None
    return bits, s


def absorbed(prefixes):
    """Sorted, deduplicated, with every prefix that a shorter kept one starts
    dropped (src/filter.rs:1326-1338)."""
    kept = []
    for p in sorted(set(prefixes)):
        if kept and p.startswith(kept[-1]):
            continue
        kept.append(p)
    return kept


def check(prefixes):
    got = filters.FilterSet.parse(prefixes).index_words()
    want = reference_index(absorbed(prefixes))
    if want is None:
        assert got is None, (len(prefixes), got and got[1])
        return None
    words, bits = got
    assert bits == want[0], (len(prefixes), bits, want[0])
    got_slots = {64 * i + b for i, w in enumerate(words) if w for b in range(64) if (w >> b) & 1}
    assert got_slots == want[1]
    return bits


def random_prefixes(rng, count, lo, hi):
    return ["".join(rng.choice(ALPHABET) for _ in range(rng.randint(lo, hi)))
            for _ in range(count)]


def test_index_width_is_the_narrowest_within_one_percent():
    rng = random.Random(21)
    # the closed form for prefixes longer than the widest width: n / 2^bits
    assert check(["abcdef"]) == 16
    assert check(random_prefixes(rng, 1000, 6, 8)) == 17     # 1000 / 2^17 = 0.76%
    assert check(random_prefixes(rng, 10000, 6, 8)) == 20    # src/filter.rs:2234-2239
    assert check(random_prefixes(rng, 700, 6, 8)) == 17      # 700 / 2^16 = 1.07%
    assert check(random_prefixes(rng, 655, 6, 8)) == 16      # 655 / 2^16 = 0.9994%


def test_short_prefixes_fill_runs_and_decide_the_width():
    rng = random.Random(22)
    # a 3-symbol prefix fills 2^(bits - 15) slots: occupancy 1/2^15 at any width
    assert check(["abc"]) == 16
    # a hundred 2-symbol prefixes never get under 1%: the widest, 9.8% full
    two = sorted({"".join(rng.choice(ALPHABET) for _ in range(2)) for _ in range(120)})[:100]
    assert check(two) == 24
    # 300 of them fill 29%: the index is discarded
    many = sorted({a + b for a in ALPHABET for b in ALPHABET})[:300]
    assert check(many) is None
    # mixed lengths that overlap
    assert check(["ab", "abcdefg", "q2", "q2x7", "zzzzzzzz"] + random_prefixes(rng, 50, 4, 9))


def test_index_over_the_dictionary_workload():
    words = workload.dictionary(10000, 2)
    assert check(words) == 20


def test_reference_cases_for_absorption_and_matching():
    # src/filter.rs:3468-3472
    assert len(filters.FilterSet.parse(["ab", "abc", "ab", "abcd", "xy"])) == 2
    # src/filter.rs:3474-3486: the prefix that matches, or none
    fs = filters.FilterSet.parse(["abc", "xyz"])
    rng = random.Random(23)
    seen = {"abc": 0, "xyz": 0, None: 0}
    for _ in range(20000):
        pk = bytes(rng.getrandbits(8) for _ in range(32))
        addr = key.address(pk)
        which = fs.match_address(key.address_bytes(pk))
        want = "abc" if addr.startswith("abc") else "xyz" if addr.startswith("xyz") else None
        assert (which >= 0) == (want is not None)
        seen[want] += 1
    assert seen[None] > 0
    # planted: the first three symbols are the top 15 bits of the key
    for text in ("abc", "xyz"):
        v, _ = prefix_value(text)
        pk = bytearray(rng.getrandbits(8) for _ in range(32))
        head = (v << 1) | (pk[1] & 1)
        pk[0], pk[1] = head >> 8, head & 0xFF
        assert key.address(bytes(pk)).startswith(text)
        assert fs.match_address(key.address_bytes(bytes(pk))) >= 0
