"""The default launch shape itself: 2^20 chains, 1024 pairs per window, one
window per launch, 2,148,532,224 candidates.

The oracle pins the small shapes bit for bit (test_gpu_parity); a CPU oracle
over 2^31 candidates would take many minutes.  So the full-size launch is
checked by shape invariance, which carries that pin to the full size: every
hit of one full-size launch must be found, and only those, by a shape of 2^18
chains and four rounds over the same candidates (for the pattern form, the
address stage then drains in place instead of in its own kernel).  A sample of
the hits is re-derived from its offset on the host.  The slot engine's default
unit shape is checked the same way against the single engine, and the sign
of x derived in the address kernel against the sign formed in the loop.
"""
import hashlib

import pytest

from onion_gen_amd import filters, key
from tests.test_gpu_parity import run_engine

pytestmark = pytest.mark.gpu

FULL = dict(threads=1 << 20, half=1024, rounds=1, hit_cap=1 << 20)
SMALL = dict(threads=1 << 18, half=256, rounds=4, hit_cap=1 << 20)
N = (1 << 20) * 2049


def say(msg):
    print(msg, flush=True)


def check(seed, line, expect_lo, expect_hi):
    start = key.start_point(seed)
    fs = filters.FilterSet.parse(line)
    say(f"case: {line}, one launch at {FULL}")
    full = run_engine(start, fs, N, **FULL)
    say(f"case: {line}, the same candidates at {SMALL}")
    small = run_engine(start, fs, N, **SMALL)
    assert expect_lo < len(full) < expect_hi
    assert full == small
    for k, packed in full[:: max(1, len(full) // 500)]:
        pk, addr = key.finish_hit(seed, 8 * k)
        assert fs.match_address(key.address_bytes(pk)) >= 0, (k, addr)
        assert pk[:31] == packed[:31] and (pk[31] & 0x7F) == packed[31]


def test_full_launch_prefix():
    # a 3-symbol prefix: 2^31 / 2^15 = 65,536 hits expected
    check(hashlib.sha256(b"full prefix").digest(), "2xe", 60000, 71000)


def test_full_launch_suffix():
    # a 5-symbol suffix: symbol 55 is always d and q is one of the four
    # possible symbols 54, so 2^31 / (2^15 * 4) = 16,384 hits expected
    check(hashlib.sha256(b"full suffix").digest(), "suffix:2xaqd", 14500, 18500)


def test_slot_engine_default_shape():
    """The slot engine at its default unit shape (64 slots x 16384 chains x
    256 pairs): a prefix search and a suffix search resident together, each
    laid out over half the slots, report over their examined prefix [0, n)
    exactly what the single engine reports over the same candidates."""
    from onion_gen_amd import batch
    cases = [(hashlib.sha256(b"slot prefix").digest(), "2xe"),
             (hashlib.sha256(b"slot suffix").digest(), "suffix:2xaqd")]
    say("case: slot engine 64 x 16384 x 256, a prefix and a suffix, 2 steps")
    got = [[] for _ in cases]
    with batch.Batch(slots=64, chains=16384, half=256, rounds=1, ring_cap=1 << 18) as b:
        ids = [b.admit(key.start_point(seed), filters.FilterSet.parse(line), i)
               for i, (seed, line) in enumerate(cases)]
        for _ in range(2):
            for h in b.step():
                got[h.tag].append((h.offset // 8, h.packed))
        examined = [b.search_state(i)["examined"] for i in ids]
    for (seed, line), hits, n in zip(cases, got, examined):
        assert n >= (1 << 29), (line, n)
        say(f"case: {line} alone over the {n} candidates the slot engine examined")
        alone = run_engine(key.start_point(seed), filters.FilterSet.parse(line), n,
                           threads=1 << 18, half=256, rounds=1, hit_cap=1 << 20)
        assert sorted(h for h in hits if h[0] < n) == alone and alone, (line, len(hits))
        assert all(h[0] < n for h in hits)


def test_full_launch_substring_with_the_sign_deferred():
    """At the default shape a rare substring leaves the sign of x to the
    address kernel, which finds each parked entry's chain from its 64-bit
    index (past 2^31 here) and derives x from that chain's window centre.  The
    launch reports exactly what the same launch reports with the sign formed
    in the search loop; a sample of the hits is re-derived on the host."""
    seed = hashlib.sha256(b"full substring").digest()
    start = key.start_point(seed)
    line = "contains:q2x7a"
    fs = filters.FilterSet.parse(line)
    from onion_gen_amd import gpu
    with gpu.Engine.with_filters(start, fs, **FULL) as eng:
        assert eng.sign_stage() == 2
    say(f"case: {line}, one launch at {FULL}, sign in the address kernel")
    deferred = run_engine(start, fs, N, **FULL)
    say(f"case: {line}, one launch at {FULL}, sign in the loop")
    in_loop = run_engine(start, fs, N, sign_stage=1, **FULL)
    # 52 placements of a 5-symbol word: about 2^31 * 52 / 2^25 = 3,328 hits
    assert 2600 < len(deferred) < 4200
    assert deferred == in_loop
    for k, packed in deferred[:: max(1, len(deferred) // 200)]:
        pk, addr = key.finish_hit(seed, 8 * k)
        assert fs.match_address(key.address_bytes(pk)) >= 0, (k, addr)
        assert pk[:31] == packed[:31] and (pk[31] & 0x7F) == packed[31]
