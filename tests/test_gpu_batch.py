"""The slot engine against the oracle: each search in a mixed batch reports
exactly its own hits, the same as it reports alone, across graph replays,
retirement and re-admission."""
import hashlib

import pytest

from onion_gen_amd import batch, filters, gpu, key
from tests.oracle_lib import OracleFilters, search

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def clear_sign(pk):
    return pk[:31] + bytes([pk[31] & 0x7F])


def oracle(seed, lines, n):
    start = key.start_point(seed)
    return sorted((k, clear_sign(pk)) for k, pk, _ in search(start, 0, n, 256, OracleFilters(lines)))


MIX = [
    (b"prefix", ["2x"]),
    (b"suffix", ["suffix:aqd"]),
    (b"dict", ["2xe", "abc", "zz7", "q2q", "77", "a2b3"]),
    (b"contains", ["contains:abcd"]),
]


def test_mixed_batch_matches_each_search_alone():
    say("case: batch 4 slots x 512 chains, half 64, mixed filter kinds")
    b = batch.Batch(slots=4, chains=512, half=64, rounds=1)
    seeds = {tag: hashlib.sha256(tag).digest() for tag, _ in MIX}
    fsets = {tag: filters.FilterSet.parse(lines) for tag, lines in MIX}
    ids = {}
    for i, (tag, lines) in enumerate(MIX):
        ids[tag] = b.admit(key.start_point(seeds[tag]), fsets[tag], i)
    assert sorted(ids.values()) == [0, 1, 2, 3]
    got = {tag: [] for tag, _ in MIX}
    n = 1 << 18
    steps = -(-n // b.per_step())
    for _ in range(steps):
        for h in b.step():
            tag = MIX[h.tag][0]
            assert h.search == ids[tag]
            if h.offset // 8 < n:
                got[tag].append((h.offset // 8, h.packed))
    for tag, lines in MIX:
        want = oracle(seeds[tag], lines, n)
        assert sorted(got[tag]) == want and want, (tag, len(got[tag]), len(want))
    # alone: the single-search engine over the same range gives the same set
    eng = gpu.Engine.with_filters(key.start_point(seeds[b"prefix"]), fsets[b"prefix"],
                                  threads=512, half=64)
    alone, done = [], 0
    while done < n:
        alone += [(s.offset // 8, s.packed) for s in eng.advance() if s.offset // 8 < n]
        done += eng.batch()
    eng.close()
    assert sorted(alone) == sorted(got[b"prefix"])
    b.close()


def test_retire_and_readmit_starts_a_search_from_zero():
    say("case: retire a search, admit a new one in its place")
    b = batch.Batch(slots=2, chains=256, half=32, rounds=1)
    s1, s2 = hashlib.sha256(b"one").digest(), hashlib.sha256(b"two").digest()
    f = filters.FilterSet.parse("2x")
    id_a = b.admit(key.start_point(s1), f, 1)
    id_b = b.admit(key.start_point(s2), f, 2)
    for _ in range(3):
        b.step()
    assert b.search_state(id_a)["examined"] == 3 * b.per_step()
    b.retire(id_a)
    assert not b.search_state(id_a)["live"]
    s3 = hashlib.sha256(b"three").digest()
    id_c = b.admit(key.start_point(s3), f, 3)
    assert id_c == id_a
    n = 1 << 16
    got = {2: [], 3: []}
    steps = -(-n // b.per_step())
    for _ in range(steps):
        for h in b.step():
            if h.tag in got and h.offset // 8 < n + 3 * b.per_step():
                got[h.tag].append((h.offset // 8, h.packed))
    want3 = oracle(s3, ["2x"], n)
    assert sorted(x for x in got[3] if x[0] < n) == want3
    # the search in the other slot simply continued: its k kept counting
    want2 = oracle(s2, ["2x"], 3 * b.per_step() + n)
    assert sorted(x for x in got[2] if x[0] >= 3 * b.per_step()) == \
        [x for x in want2 if x[0] >= 3 * b.per_step()]
    b.close()


def test_more_searches_than_slots_share_the_device_round_robin():
    """Six resident searches on two slots: each step runs two of them, every
    search advances in turn, and each one's hits over what it examined are
    exactly its oracle's."""
    say("case: 6 resident searches on 2 slots x 256 chains, half 32")
    b = batch.Batch(slots=2, chains=256, half=32, rounds=1)
    f = filters.FilterSet.parse(["2x", "suffix:aqd"])
    seeds = [hashlib.sha256(b"rr%d" % i).digest() for i in range(6)]
    ids = [b.admit(key.start_point(s), f, i) for i, s in enumerate(seeds)]
    got = {i: [] for i in range(6)}
    for _ in range(9):  # three turns each
        for h in b.step():
            got[h.tag].append((h.offset // 8, h.packed))
    for i, sid in enumerate(ids):
        st = b.search_state(sid)
        assert st["examined"] == 3 * b.per_step()
        want = oracle(seeds[i], ["2x", "suffix:aqd"], st["examined"])
        assert sorted(got[i]) == want and want, i
    b.close()


def test_hit_quota_records_only_what_the_search_wants():
    """A search that wants 5 hits records 5 (real ones, each among the
    oracle's); every hit is still counted; the search beside it, with no
    quota, records all of its own."""
    say("case: quota 5 on an easy filter, 2 slots x 256 chains, half 32")
    b = batch.Batch(slots=2, chains=256, half=32, rounds=1)
    s1, s2 = hashlib.sha256(b"quota").digest(), hashlib.sha256(b"free").digest()
    f = filters.FilterSet.parse("suffix:aqd")
    a = b.admit(key.start_point(s1), f, 1)
    c = b.admit(key.start_point(s2), f, 2)
    b.set_want(a, 5)
    got = {1: [], 2: []}
    for _ in range(4):
        for h in b.step():
            got[h.tag].append((h.offset // 8, h.packed))
    n = b.search_state(a)["examined"]
    want1 = oracle(s1, ["suffix:aqd"], n)
    want2 = oracle(s2, ["suffix:aqd"], n)
    assert len(want1) > 5
    assert len(got[1]) == 5 and set(got[1]) <= set(want1)
    assert b.found(a) == len(want1) and b.search_state(a)["hits"] == 5
    assert sorted(got[2]) == want2 and b.found(c) == len(want2)
    b.close()


def test_searches_widen_into_idle_slots_and_narrow_under_load():
    """Widths follow the resident count: a lone search takes every unit slot,
    new arrivals narrow it, retirements widen the rest again.  Each relayout
    starts from the search's next candidate, so every search's hits over what
    it examined are exactly the oracle's."""
    say("case: 8 unit slots x 256 chains, half 32; 2 -> 8 -> 2 resident searches")
    b = batch.Batch(slots=8, chains=256, half=32, rounds=1)
    f = filters.FilterSet.parse(["2x"])
    seeds = [hashlib.sha256(b"wide%d" % i).digest() for i in range(8)]
    got = {i: [] for i in range(8)}

    def steps(n):
        for _ in range(n):
            for h in b.step():
                got[h.tag].append((h.offset // 8, h.packed))

    ids = [b.admit(key.start_point(seeds[0]), f, 0)]
    assert b.width(ids[0]) == 8
    steps(1)
    ids.append(b.admit(key.start_point(seeds[1]), f, 1))
    steps(2)
    assert [b.width(i) for i in ids] == [4, 4]
    ids += [b.admit(key.start_point(seeds[i]), f, i) for i in range(2, 8)]
    steps(3)
    assert [b.width(i) for i in ids] == [1] * 8
    for i in ids[2:]:
        b.retire(i)
    steps(2)
    assert [b.width(i) for i in ids[:2]] == [4, 4]
    total = 0
    for i, sid in enumerate(ids):
        n = b.search_state(sid)["examined"]
        total += n
        assert sorted(got[i]) == oracle(seeds[i], ["2x"], n), i
    assert total == b.examined()
    b.close()


def test_two_round_steps_drain_the_address_stage_in_place():
    """A step of one round leaves the pattern family's address stage to its own
    kernel; a step of two rounds drains it inside the search kernel.  Both
    report the oracle's hits."""
    say("case: batch steps of 2 rounds, suffix and contains searches")
    mix = [(b"s2", ["suffix:aqd"]), (b"c2", ["contains:abcd"])]
    for rounds in (1, 2):
        b = batch.Batch(slots=2, chains=256, half=32, rounds=rounds)
        ids = {tag: b.admit(key.start_point(hashlib.sha256(tag).digest()),
                            filters.FilterSet.parse(lines), i)
               for i, (tag, lines) in enumerate(mix)}
        n = 1 << 16
        got = {tag: [] for tag, _ in mix}
        for _ in range(-(-n // b.per_step())):
            for h in b.step():
                tag = mix[h.tag][0]
                assert h.search == ids[tag]
                if h.offset // 8 < n:
                    got[tag].append((h.offset // 8, h.packed))
        b.close()
        for tag, lines in mix:
            want = oracle(hashlib.sha256(tag).digest(), lines, n)
            assert sorted(got[tag]) == want and want, (rounds, tag, len(got[tag]), len(want))
