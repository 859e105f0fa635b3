"""The regex: form on the device (src/expression.rs, src/filter.rs:884-1005)
against the expression applied to every candidate's finished address.

The oracle enumerates every candidate of the range (a set of all 32 one-symbol
prefixes accepts every address) with its public key and address; Python's `re`
decides each address.  The device must report exactly those candidates: its
proxies may only narrow what the automata see, never what they accept.
"""
import hashlib
import re

import pytest

from onion_gen_amd import batch, filters, gpu, key, workload
from tests.oracle_lib import OracleFilters, search

pytestmark = pytest.mark.gpu

EVERY = list(workload.ALPHABET)  # one-symbol prefixes: every address matches one


def say(msg):
    print(msg, flush=True)


def clear_sign(pk):
    return pk[:31] + bytes([pk[31] & 0x7F])


def truth(start, expressions, n):
    """(k, packed) of every k < n whose address one of the expressions matches."""
    rx = [re.compile(e) for e in expressions]
    every = search(start, 0, n, 256, OracleFilters(EVERY), cap=n)
    assert len(every) == n
    return sorted((k, clear_sign(pk)) for k, pk, addr in every
                  if any(r.search(addr[:56]) for r in rx))


def run_engine(start, fs, n, **shape):
    eng = gpu.Engine.with_filters(start, fs, **shape)
    got, done = [], 0
    try:
        while done < n:
            got += [(s.offset // 8, s.packed) for s in eng.advance() if s.offset // 8 < n]
            done += eng.batch()
    finally:
        eng.close()
    return sorted(got)


def filler_contains(seed):
    """Eight 32-symbol substrings that fill all 256 automaton-bank bits, so an
    expression's proxies find no room and its automaton runs on every
    candidate.  Random, so none matches in the ranges tested."""
    return ["contains:" + "".join(workload.ALPHABET[b & 31]
                                  for b in hashlib.sha256(seed + bytes([i])).digest())
            for i in range(8)]


CASES = [
    # (filter lines, the Python expressions that decide, n)
    (["regex:^ab[2-7]"], ["^ab[2-7]"], 1 << 18),            # literal prefixes in the key tables
    (["regex:^a[2-7]b", "regex:^(cat|dog)"], ["^a[2-7]b", "^(cat|dog)"], 1 << 18),
    (["regex:zz[2-7]z"], ["zz[2-7]z"], 1 << 18),            # substring proxy into the tail
    (["regex:^zz.*qd$"], ["^zz.*qd$"], 1 << 18),            # decided on the checksum symbols
    (["regex:^[bcdfghjklmnp]{4}"], ["^[bcdfghjklmnp]{4}"], 1 << 16),  # class proxy, many hits
    (["regex:(cat|dog)food|^my(shop|store)|q[2-7]{2}q"],
     ["(cat|dog)food|^my(shop|store)|q[2-7]{2}q"], 1 << 17),
    (["regex:^ab[2-7]", "regex:[2-7]{3}qd$"], ["^ab[2-7]", "[2-7]{3}qd$"], 1 << 17),
    (["zz", "regex:^ab[2-7]", "suffix:[a-e]id"], ["^zz", "^ab[2-7]", "[a-e]id$"], 1 << 16),
    (filler_contains(b"every") + ["regex:zz[2-7]z"], ["zz[2-7]z"], 1 << 17),  # no proxy
    (["regex:\\Bq[2-7]\\B"], ["(?<=.)q[2-7](?=.)"], 1 << 16),  # looks in the middle
]


@pytest.mark.parametrize("lines,python,n", CASES)
def test_engine_reports_exactly_the_matching_candidates(lines, python, n):
    say(f"case: regex engine {lines[-1]} (+{len(lines) - 1}) n={n}")
    start = key.start_point(hashlib.sha256(repr(lines).encode()).digest())
    fs = filters.FilterSet.parse(lines)
    got = run_engine(start, fs, n, threads=1024, half=128)
    want = truth(start, python, n)
    assert want, "the range holds no match: the case proves nothing"
    assert got == want, (len(got), len(want))


def test_each_case_takes_the_path_it_is_meant_to():
    def lines_of(lines):
        return filters.FilterSet.parse(lines).describe()
    assert any("key tables" in l for l in lines_of(["regex:^ab[2-7]"]))
    assert any("key tables" in l for l in lines_of(["regex:^a[2-7]b", "regex:^(cat|dog)"]))
    assert any("automaton banks" in l for l in lines_of(["regex:^[bcdfghjklmnp]{4}"]))
    assert any("every candidate" in l
               for l in lines_of(filler_contains(b"every") + ["regex:zz[2-7]z"]))


def test_batch_runs_expressions_beside_other_searches():
    say("case: batch 4 slots, regex searches beside a prefix search")
    mix = [(b"r1", ["regex:zz[2-7]z"], ["zz[2-7]z"]),
           (b"r2", ["regex:^zz.*qd$"], ["^zz.*qd$"]),
           (b"p", ["2x"], ["^2x"]),
           (b"r3", ["regex:^a[2-7]c"], ["^a[2-7]c"])]
    b = batch.Batch(slots=4, chains=512, half=64, rounds=1)
    ids = {}
    for i, (tag, lines, _) in enumerate(mix):
        ids[tag] = b.admit(key.start_point(hashlib.sha256(tag).digest()),
                           filters.FilterSet.parse(lines), i)
    n = 1 << 17
    got = {tag: [] for tag, _, _ in mix}
    for _ in range(-(-n // b.per_step())):
        for h in b.step():
            tag = mix[h.tag][0]
            assert h.search == ids[tag]
            if h.offset // 8 < n:
                got[tag].append((h.offset // 8, h.packed))
    b.close()
    for tag, _, python in mix:
        want = truth(key.start_point(hashlib.sha256(tag).digest()), python, n)
        assert sorted(got[tag]) == want and want, (tag, len(got[tag]), len(want))


def test_dictionary_of_general_forms_matches_oracle():
    """General forms beyond the automaton banks (a dictionary of substrings,
    suffixes and class prefixes) run as one automaton on every candidate; the
    engine's survivors equal the oracle's hits.  A tenth of the dictionary is
    taken from addresses in the range, so every kind hits."""
    import random
    say("case: dictionary automaton, 1,500 general forms, n=2^17")
    n = 1 << 17
    start = key.start_point(hashlib.sha256(b"dictionary").digest())
    every = search(start, 0, n, 256, OracleFilters(EVERY), cap=n)
    rng = random.Random(9)
    A = workload.ALPHABET
    lines = ["contains:" + "".join(rng.choice(A) for _ in range(8)) for _ in range(1200)]
    lines += ["contains:" + "".join(rng.choice(A) for _ in range(5)) + "?" for _ in range(100)]
    lines += ["suffix:" + "".join(rng.choice(A) for _ in range(6)) + rng.choice("aiqy") + "d"
              for _ in range(60)]
    for _, _, addr in rng.sample(every, 150):
        kind = rng.randrange(3)
        if kind == 0:
            s = rng.randint(0, 48)
            lines.append("contains:" + addr[s:s + 8])
        elif kind == 1:
            lines.append("suffix:" + addr[56 - rng.randint(7, 10):56])
        else:
            lines.append("[" + addr[0] + "7]" + addr[1:7])
    fs = filters.FilterSet.parse(lines)
    assert any(l.startswith("dictionary automaton") for l in fs.describe())
    got = run_engine(start, fs, n, threads=1024, half=64)
    want = sorted((k, clear_sign(pk)) for k, pk, _ in search(start, 0, n, 256, OracleFilters(lines)))
    assert len(want) >= 150
    assert got == want, (len(got), len(want))


def test_dictionary_split_across_automata_matches_every_address():
    """10,000 eight-symbol substrings exceed one automaton's states and are
    split across two, behind the window prefilter (about 11% of keys go to the
    address stage through a window hit); the engine reports exactly the
    candidates whose address holds one of the words (decided here by a set of
    the words over every 8-symbol window of the oracle's addresses)."""
    import random
    say("case: 10,000 substrings in two automata, n=2^16")
    n = 1 << 16
    start = key.start_point(hashlib.sha256(b"split dictionary").digest())
    every = search(start, 0, n, 256, OracleFilters(EVERY), cap=n)
    rng = random.Random(10)
    A = workload.ALPHABET
    words = {"".join(rng.choice(A) for _ in range(8)) for _ in range(10000)}
    for _, _, addr in rng.sample(every, 60):
        s = rng.randint(0, 48)
        words.add(addr[s:s + 8])
    fs = filters.FilterSet.parse(["contains:" + w for w in sorted(words)])
    assert any(l.startswith("dictionary automata: 2 of") for l in fs.describe())
    assert any(l.startswith("window prefilter") for l in fs.describe())
    got = run_engine(start, fs, n, threads=1024, half=64)
    want = sorted((k, clear_sign(pk)) for k, pk, addr in every
                  if any(addr[s:s + 8] in words for s in range(49)))
    assert len(want) >= 60
    assert got == want, (len(got), len(want))


def test_dictionary_behind_the_window_prefilter_matches_every_address():
    """1,200 eight-symbol substrings (one automaton in global memory) take the
    window prefilter: every key walks symbols 42..48, and a key whose 7-symbol
    window hits both maps goes to the address stage.  A twentieth of the words
    are taken from the range's addresses at every placement, those that run
    into the checksum included.  The engine reports exactly the candidates
    whose address holds a word, as a set of the words over every 8-symbol
    window of the oracle's addresses decides, in one-round launches (the
    address kernel) and in two-round launches (drained in place)."""
    import random
    say("case: 1,200 substrings behind the window prefilter, n=2^16")
    n = 1 << 16
    start = key.start_point(hashlib.sha256(b"prefiltered dictionary").digest())
    every = search(start, 0, n, 256, OracleFilters(EVERY), cap=n)
    rng = random.Random(12)
    A = workload.ALPHABET
    words = {"".join(rng.choice(A) for _ in range(8)) for _ in range(1140)}
    for i, (_, _, addr) in enumerate(rng.sample(every, 60)):
        s = i % 49
        words.add(addr[s:s + 8])
    fs = filters.FilterSet.parse(["contains:" + w for w in sorted(words)])
    assert any(l.startswith("window prefilter") for l in fs.describe())
    want = sorted((k, clear_sign(pk)) for k, pk, addr in every
                  if any(addr[s:s + 8] in words for s in range(49)))
    assert len(want) >= 60
    for rounds in (1, 2):
        got = run_engine(start, fs, n, threads=1024, half=64, rounds=rounds)
        assert got == want, (rounds, len(got), len(want))


def test_dictionary_that_reaches_the_checksum_matches_every_address():
    """1,000 five-symbol substrings: a word can still end in the checksum for
    nearly every candidate, so the key stage is skipped and every candidate
    parks; each lane drains a share of many entries.  The engine reports
    exactly the candidates whose address holds a word anywhere in its 56
    symbols, in one-round launches (the address kernel) and in two-round
    launches (drained in place), alone and beside a literal prefix, whose
    hits the key stage still decides."""
    import random
    say("case: 1,000 five-symbol substrings, n=2^17")
    n = 1 << 17
    start = key.start_point(hashlib.sha256(b"checksum dictionary").digest())
    every = search(start, 0, n, 256, OracleFilters(EVERY), cap=n)
    rng = random.Random(11)
    A = workload.ALPHABET
    words = {"".join(rng.choice(A) for _ in range(5)) for _ in range(1000)}
    lines = ["contains:" + w for w in sorted(words)]
    fs = filters.FilterSet.parse(lines)
    assert any(l.startswith("key stage skipped") for l in fs.describe())
    want = sorted((k, clear_sign(pk)) for k, pk, addr in every
                  if any(addr[s:s + 5] in words for s in range(52)))
    assert len(want) > 100
    for rounds in (1, 2):
        got = run_engine(start, fs, n, threads=256, half=64, rounds=rounds)
        assert got == want, (rounds, len(got), len(want))
    say("case: the same beside the prefix zz, n=2^17")
    fs = filters.FilterSet.parse(lines + ["zz"])
    assert any(l.startswith("key stage skipped") for l in fs.describe())
    want = sorted((k, clear_sign(pk)) for k, pk, addr in every
                  if addr.startswith("zz") or any(addr[s:s + 5] in words for s in range(52)))
    assert sum(1 for k, pk, addr in every if addr.startswith("zz")) > 50
    got = run_engine(start, fs, n, threads=256, half=64)
    assert got == want, (len(got), len(want))


def test_address_stage_in_place_and_deferred_agree():
    """A one-round launch leaves the address stage (sign, SHA3, exact pass) to
    its own kernel; a launch of several rounds drains it in place.  Over the
    same candidates both report the same survivors, and those equal the
    oracle's."""
    say("case: address stage deferred (rounds 1) vs in place (rounds 2)")
    lines = ["suffix:[a-e]aqd", "contains:zz7", "regex:^zz.*qd$"]
    n = 1 << 17
    start = key.start_point(hashlib.sha256(b"address stage").digest())
    fs = filters.FilterSet.parse(lines)
    deferred = run_engine(start, fs, n, threads=512, half=64, rounds=1)
    in_place = run_engine(start, fs, n, threads=512, half=64, rounds=2)
    want = truth(start, ["[a-e]aqd$", "zz7", "^zz.*qd$"], n)
    assert want and deferred == want and in_place == want, (len(deferred), len(in_place), len(want))
