"""Parity of the HIP path with the CPU oracle, through the C ABI.

Bit-exact: the set of (k, packed key) the device reports must equal the
oracle's hits for the same start scalar and filters (src/batch.rs,
src/filter.rs, src/run.rs:1030-1116 semantics).  Every test prints the case
it runs before touching the device.
"""
import ctypes
import hashlib
import json
import os
import random

import pytest

from onion_gen_amd import _lib, filters, gpu, key, workload
from tests import ed25519_ref as ref
from tests.oracle_lib import OracleFilters, digest, load, search

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = ref.P


def say(msg):
    print(msg, flush=True)


def clear_sign(pk):
    return pk[:31] + bytes([pk[31] & 0x7F])


def run_engine(start, fs, n, **shape):
    """All survivors with k < n, sorted (k, packed)."""
    eng = gpu.Engine.with_filters(start, fs, **shape)
    got, done = [], 0
    try:
        while done < n:
            for s in eng.advance():
                assert s.offset % 8 == 0
                if s.offset // 8 < n:
                    got.append((s.offset // 8, s.packed))
            done += eng.batch()
    finally:
        eng.close()
    return sorted(got)


def oracle_hits(start, lines, n, width=256):
    return [(k, clear_sign(pk)) for k, pk, _ in search(start, 0, n, width, OracleFilters(lines))]


def field_op(op, a, b):
    n = len(a)
    out = ctypes.create_string_buffer(32 * n)
    e = _lib.err_buf()
    rc = _lib.lib().og_hip_field_op(0, op, b"".join(x.to_bytes(32, "little") for x in a),
                                    b"".join(x.to_bytes(32, "little") for x in b), out, n, e, len(e))
    assert rc == 0, e.value
    return [int.from_bytes(out.raw[32 * i:32 * i + 32], "little") for i in range(n)]


def test_device_is_visible():
    say("case: device discovery")
    devs = gpu.look()
    assert devs and "gfx950" in devs[0].name


def test_field_kernels_match_big_integers():
    say("case: field kernels, 20000 random + edge values")
    rng = random.Random(5)
    edges = [0, 1, 2, 19, P - 1, P, P + 1, 2 * P, 2 ** 255 - 1, 2 ** 255, 2 ** 256 - 1,
             2 ** 256 - 38, 2 ** 256 - 39, 38]
    a = edges + [rng.getrandbits(256) for _ in range(20000)]
    b = list(reversed(edges)) + [rng.getrandbits(256) for _ in range(20000)]
    assert field_op(0, a, b) == [x * y % P for x, y in zip(a, b)]
    assert field_op(1, a, b) == [x * x % P for x in a]
    assert field_op(2, a, b) == [(x + y) % P for x, y in zip(a, b)]
    assert field_op(3, a, b) == [(x - y) % P for x, y in zip(a, b)]
    assert field_op(5, a, b) == [x % P for x in a]
    nz = [x for x in a if x % P][:2000]
    assert field_op(4, nz, nz) == [pow(x, P - 2, P) for x in nz]
    # the 10 x 25.5 form the search kernels run in
    assert field_op(10, a, b) == [x * y % P for x, y in zip(a, b)]
    assert field_op(11, a, b) == [x * x % P for x in a]
    assert field_op(12, a, b) == [(x + y) % P for x, y in zip(a, b)]
    assert field_op(13, a, b) == [(x - y) % P for x, y in zip(a, b)]
    assert field_op(15, a, b) == [x % P for x in a]
    assert field_op(14, nz, nz) == [pow(x, P - 2, P) for x in nz]
    assert field_op(16, a, b) == [(x - y) * (2 * x - y) % P for x, y in zip(a, b)]
    assert field_op(25, a, b) == [(x - y) * y % P for x, y in zip(a, b)]
    # the divstep inversion, against Fermat, over the edges and random values
    # (0 maps to 0), and on values in [p, 2^255) as a product leaves them
    inv_in = a[:2017] + [P, P + 1, P + 18, 2 ** 255 - 1, 2 ** 255 + P, 2 ** 256 - 1]
    assert field_op(26, inv_in, inv_in) == [pow(x % P, P - 2, P) for x in inv_in]
    # a product's output canonicalised as the search does it (no carry pass
    # first): values in [p, 2^255) kept as they stand, and products
    above = [P, P + 1, P + 18, 2 ** 255 - 1, 2 ** 255 + P, 2 ** 256 - 1]
    assert field_op(17, a + above, a + above) == [(x % 2 ** 255) % P for x in a + above]
    assert field_op(18, a, b) == [x * y % P for x, y in zip(a, b)]
    assert field_op(19, a, b) == [(x % 2 ** 255) * y % P & 1 for x, y in zip(a, b)]
    assert field_op(19, above, [1] * len(above)) == [(x % 2 ** 255) % P & 1 for x in above]
    # products by a constant as the search loop makes them from its table (the
    # constant is the operand of element i & ~63, uniform over a wave): one
    # product, and a double product with one carry pass
    ra = [rng.getrandbits(256) for _ in range(4096)] + a[:1000]
    rb = [rng.getrandbits(256) for _ in range(4096)] + b[:1000]
    lead_a = [ra[i & ~63] for i in range(len(ra))]
    lead_b = [rb[i & ~63] for i in range(len(rb))]
    assert field_op(21, ra, rb) == [x * c % P for x, c in zip(ra, lead_b)]
    assert field_op(22, ra, rb) == [(x * c + y * e) % P
                                    for x, y, c, e in zip(ra, rb, lead_b, lead_a)]
    # the key forms' probe word (low 32 bits) and its exactness flag (bit 32):
    # where exact it is the canonical low word of (2x - y)(x - y), else that
    # minus it is 0, 19 or 38; rarely inexact
    got = field_op(23, a, b)
    diffs = [((2 * x - y) * (x - y) % P - w) % 2 ** 32 for x, y, w in zip(a, b, got)]
    exact = [w >> 32 == 1 for w in got]
    assert all(d == 0 for d, e in zip(diffs, exact) if e)
    assert all(d in (0, 19, 38) for d in diffs)
    assert exact[len(edges):].count(False) <= len(a) // 200  # random operands: about 1 in 1,024
    # the three-column word (op 27): bits 0..25 of the same, on the same terms,
    # and op 23's word less column 1
    got26 = field_op(27, a, b)
    d26 = [((2 * x - y) * (x - y) % P - w) % 2 ** 26 for x, y, w in zip(a, b, got26)]
    assert [w >> 32 == 1 for w in got26] == exact
    assert all(d == 0 for d, e in zip(d26, exact) if e)
    assert all(d in (0, 19, 38) for d in d26)
    assert [w & (2 ** 26 - 1) for w in got26] == [w % 2 ** 26 for w in got]
    assert all(w % 2 ** 32 < 2 ** 26 for w in got26)
    # the sign of x as the search takes it: from three columns where they settle it
    assert field_op(24, a, b) == [(x % 2 ** 255) * y % P & 1 for x, y in zip(a, b)]
    assert field_op(24, above, [1] * len(above)) == [(x % 2 ** 255) % P & 1 for x in above]


def test_device_checksum_matches_sha3():
    """The gfx950 Keccak (32-bit halves, alignbit rotations, bitop3 chi, cut
    last round) against hashlib's SHA3-256 on random keys, sign bits set and
    clear, and the all-zero and all-one keys."""
    say("case: device address checksum, 20000 random keys")
    rng = random.Random(11)
    keys = [0, 2 ** 256 - 1, 2 ** 255, 2 ** 255 - 1] + [rng.getrandbits(256) for _ in range(20000)]
    got = field_op(20, keys, keys)
    for k, g in zip(keys, got):
        pk = k.to_bytes(32, "little")
        d = hashlib.sha3_256(b".onion checksum" + pk + b"\x03").digest()
        assert g == d[0] | (d[1] << 8), pk.hex()


def test_config1_golden_hits():
    """The CPU-sized workload on the device: 2^24 candidates, the 4-symbol
    prefix drawn from ChaCha20 seeded 1, the committed hit list."""
    with open(os.path.join(GOLDEN, "config1_hits.json")) as f:
        g = json.load(f)
    say(f"case: config1 prefix={g['prefix']} n=2^24")
    seed = bytes.fromhex(g["seed"])
    start = key.start_point(seed)
    got = run_engine(start, filters.FilterSet.parse(g["prefix"]), g["candidates"],
                     threads=1 << 15, half=256)
    want = [(h["k"], clear_sign(bytes.fromhex(h["pk"]))) for h in g["hits"]]
    assert got == want
    for k, _ in got:
        pk, addr = key.finish_hit(seed, 8 * k)
        assert addr.startswith(g["prefix"])


@pytest.mark.parametrize("prefix,n,shape", [
    ("2x", 1 << 18, dict(threads=256, half=64)),
    ("2xe", 1 << 21, dict(threads=1024, half=128, rounds=2)),
    ("q", 1 << 15, dict(threads=256, half=32, hit_cap=1 << 12)),
])
def test_prefix_hits_match_oracle(prefix, n, shape):
    say(f"case: prefix {prefix} n={n} shape={shape}")
    seed = hashlib.sha256(prefix.encode()).digest()
    start = key.start_point(seed)
    assert run_engine(start, filters.FilterSet.parse(prefix), n, **shape) == \
        oracle_hits(start, prefix, n)


@pytest.mark.parametrize("k,length", [(1000, 5), (50000, 6), (200000, 7)])
def test_long_prefix_taken_from_a_candidate(k, length):
    """Prefixes of 5 to 7 symbols reach past the probe's three key bytes: the
    probe passes on its 24 bits and the whole product settles the rest.  The
    prefix is the address of candidate k (from the oracle's chain), so the
    search has at least that hit."""
    from tests.oracle_lib import address, batch_chain
    start = key.start_point(hashlib.sha256(b"long prefix %d" % k).digest())
    n = 1 << 18
    pk = batch_chain(start, 256, n // 256)[k]
    prefix = address(pk)[:length]
    say(f"case: prefix {prefix} from candidate {k}, n=2^18")
    got = run_engine(start, filters.FilterSet.parse(prefix), n, threads=512, half=128)
    want = oracle_hits(start, prefix, n)
    assert got == want and k in [h[0] for h in want]


def test_shape_does_not_change_results():
    """Chain count, window width and rounds are performance knobs only
    (src/batch.rs:235-264)."""
    say("case: shape invariance, prefix 2x, n=2^17")
    start = key.start_point(bytes(range(32)))
    fs = filters.FilterSet.parse("2x")
    base = run_engine(start, fs, 1 << 17, threads=256, half=16)
    for shape in (dict(threads=512, half=100), dict(threads=256, half=511, rounds=3)):
        assert run_engine(start, fs, 1 << 17, **shape) == base


@pytest.mark.parametrize("index_bits", [24, 0, 28])
def test_dictionary_bitmap_and_exact_table(index_bits):
    """The device's map in front of the exact table: 24 bits when the set's
    own index is 24 bits or narrower (0: the automatic choice, 20 bits for
    10,000 prefixes), the set's index when wider (28: the probe word then
    takes its fourth column)."""
    say(f"case: dictionary 10000 prefixes 5..8, index bits {index_bits}, n=2^22")
    words = workload.dictionary(10000, seed=2)
    fs = filters.FilterSet.parse(words, index_bits=index_bits)
    assert fs.index_words()[1] == (index_bits or 20)
    start = key.start_point(workload.chacha_stream(3, 32))
    n = 1 << 22
    got = run_engine(start, fs, n, threads=1 << 13, half=256)
    want = oracle_hits(start, words, n)
    assert got == want and len(want) > 0


def test_dictionary_of_mixed_lengths():
    """A dictionary whose prefixes run from 1 to 13 symbols: the device's 24-bit
    map holds runs for the short ones, and the exact table one bucket and
    directory per length (the 13-symbol one compared on its first 64 bits)."""
    say("case: dictionary of 1- to 13-symbol prefixes, n=2^18")
    rng = random.Random(31)
    words = ["a", "2x", "q7z", "bcdefghijklmn"]
    words += ["".join(rng.choice(workload.ALPHABET) for _ in range(rng.randint(4, 9)))
              for _ in range(300)]
    start = key.start_point(workload.chacha_stream(5, 32))
    n = 1 << 18
    got = run_engine(start, filters.FilterSet.parse(words), n, threads=1024, half=128)
    want = oracle_hits(start, words, n)
    assert got == want and len(want) > 8000


@pytest.mark.parametrize("lines", [["contains:q2x", "contains:7ab"], ["contains:zz"],
                                   ["contains:q2x", "suffix:aqd"]])
def test_literal_banks_match_oracle(lines):
    """Sets whose key pass compares literals at each placement (one or two
    substrings, and a substring beside a suffix that only the tail sees)
    report exactly the oracle's hits."""
    say(f"case: literal bank {lines}, n=2^18")
    fs = filters.FilterSet.parse(lines)
    assert any(l.startswith("literal compare: 1") for l in fs.describe())
    start = key.start_point(hashlib.sha256(" ".join(lines).encode()).digest())
    n = 1 << 18
    got = run_engine(start, fs, n, threads=1024, half=128)
    want = oracle_hits(start, lines, n)
    assert got == want and len(want) > 100


def test_reference_constructor_bitmap_only():
    """Engine::new with the bitmap alone over-reports but never under-reports."""
    say("case: bitmap-only engine, 4 prefixes, n=2^20")
    lines = ["2xe", "abc", "zz7", "q2q"]
    fs = filters.FilterSet.parse(lines)
    words, bits = fs.index_words()
    start = key.start_point(bytes(range(5, 37)))
    eng = gpu.Engine.new(start, 1024, words, bits)
    got, done, n = set(), 0, 1 << 20
    while done < n:
        got |= {(s.offset // 8, s.packed) for s in eng.advance() if s.offset // 8 < n}
        done += eng.batch()
    eng.close()
    want = set(oracle_hits(start, lines, n))
    assert want <= got
    for k, packed in got:
        slot = int.from_bytes(packed[:8], "big") >> (64 - bits)
        assert (words[slot >> 6] >> (slot & 63)) & 1


@pytest.mark.parametrize("lines,n", [
    (["suffix:aqd"], 1 << 16),
    (["contains:2xeq"], 1 << 16),
    (["suffix:[a-e]id", "contains:zz7", "ab"], 1 << 15),
])
def test_checksum_forms_match_oracle(lines, n):
    say(f"case: checksum forms {lines} n={n}")
    start = key.start_point(hashlib.sha256(repr(lines).encode()).digest())
    got = run_engine(start, filters.FilterSet.parse(lines), n, threads=256, half=64)
    want = oracle_hits(start, lines, n)
    assert got == want and len(want) > 0


def test_overflow_is_an_error():
    say("case: survivor overflow, prefix 'a', hit_cap 4096")
    start = key.start_point(bytes(32))
    eng = gpu.Engine.with_filters(start, filters.FilterSet.parse("a"), threads=1024, half=256,
                                  hit_cap=4096)
    with pytest.raises(_lib.OgError, match="survivors"):
        eng.advance()
    eng.close()


def test_digest_over_2p20_matches_oracle():
    """A non-vacuous check of every candidate: a 1-symbol prefix hits 1/32 of
    the keys; their XOR and sum over 2^20 candidates equal the oracle's over the
    same keys."""
    say("case: digest, prefix '2', n=2^20")
    start = key.start_point(workload.config1()["seed"])
    got = run_engine(start, filters.FilterSet.parse("2"), 1 << 20, threads=1024, half=256,
                     hit_cap=1 << 16)
    want = oracle_hits(start, "2", 1 << 20)
    assert got == want


@pytest.mark.parametrize("lines,python,n", [
    (["contains:2xeq"], None, 1 << 16),
    (["suffix:aqd"], None, 1 << 15),
    (["suffix:[a-e]id", "contains:zz7", "ab"], None, 1 << 15),
    (["regex:(cat|dog)f"], ["(cat|dog)f"], 1 << 16),
])
def test_sign_of_x_in_the_loop_or_the_address_kernel(lines, python, n):
    """The sign of x formed in the search loop (stage 1) or derived in the
    address kernel from the window's centre, with one inversion per chunk of
    parked entries (stage 2): the same hits as the oracle either way (for an
    expression, the oracle's addresses decided by Python's `re`).  Parking is
    compacted over a wave, so a lane derives signs for entries of its
    neighbours' chains.  At half 64 a chunk holds 60 entries, so the suffix
    sets span several chunks."""
    start = key.start_point(hashlib.sha256(b"sign" + repr(lines).encode()).digest())
    if python:
        from tests.test_gpu_regex import truth
        want = truth(start, python, n)
    else:
        want = oracle_hits(start, lines, n)
    assert want
    for stage in (1, 2):
        say(f"case: sign stage {stage}, {lines} n={n}")
        fs = filters.FilterSet.parse(lines)
        with gpu.Engine.with_filters(start, fs, threads=256, half=64, sign_stage=stage) as eng:
            assert eng.sign_stage() == stage
        assert run_engine(start, fs, n, threads=256, half=64, sign_stage=stage) == want


def test_sign_stage_chosen_from_the_filters():
    say("case: automatic sign stage: a substring defers, a short suffix does not")
    start = key.start_point(bytes(32))
    with gpu.Engine.with_filters(start, filters.FilterSet.parse("contains:q2x7a"),
                                 threads=256, half=64) as eng:
        assert eng.sign_stage() == 2
    with gpu.Engine.with_filters(start, filters.FilterSet.parse("suffix:2xaqd"),
                                 threads=256, half=64) as eng:
        assert eng.sign_stage() == 1
    # two rounds per launch drain in place, so the loop forms the sign
    with gpu.Engine.with_filters(start, filters.FilterSet.parse("contains:q2x7a"),
                                 threads=256, half=64, rounds=2, sign_stage=2) as eng:
        assert eng.sign_stage() == 1


@pytest.mark.parametrize("bits", [16, 20, 24, 28])
def test_bitmap_survivors_are_exactly_the_slots_set(bits):
    """The bitmap probe through each workgroup's LDS copy: the whole map at
    16 bits, its 18-bit or-fold in front of the map in L2 at 20, 24 and 28
    bits (28: the probe word takes its fourth column).
    A bitmap-only engine (Engine::new, src/gpu/mod.rs:346-353) must report
    exactly the candidates whose canonical key's leading `bits` bits name a set
    slot - every key of the range enumerated by the oracle's BatchEngine."""
    from tests.oracle_lib import batch_chain
    lines = workload.dictionary(10000, seed=2)
    words, got_bits = filters.FilterSet.parse(lines, index_bits=bits).index_words()
    assert got_bits == bits
    start = key.start_point(workload.chacha_stream(4, 32))
    n = 1 << 16
    say(f"case: bitmap-only engine, 10,000 prefixes at {bits} bits, n=2^16")
    eng = gpu.Engine.new(start, 1024, words, bits)
    assert eng.batch() >= n
    got = sorted((s.offset // 8, s.packed) for s in eng.advance() if s.offset // 8 < n)
    eng.close()
    want = []
    for k, pk in enumerate(batch_chain(start, 256, n // 256)):
        packed = pk[:31] + bytes([pk[31] & 0x7F])
        slot = int.from_bytes(packed[:8], "big") >> (64 - bits)
        if (words[slot >> 6] >> (slot & 63)) & 1:
            want.append((k, packed))
    assert got == want and want, (len(got), len(want))
