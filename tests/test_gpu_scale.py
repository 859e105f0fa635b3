"""The bench's own launch shape against the oracle, at the scale it runs.

One launch of the default shape is 2^20 chains x 1024 pairs per window x one
window: 2,148,532,224 candidates.  The golden fixtures (tests/golden/
make_golden.py, the CPU oracle's BatchEngine over the same seed) hold the hits
of windows that launch covers, so the full-size launch is compared with the
oracle itself, not with another shape of the same kernel:
  - the single-GPU workload's seed with the first 4 symbols of its 6-symbol
    prefix over k < 2^28: every hit, its packed key, and its public key with
    the sign as the hit sink re-derives it;
  - the first 3 symbols over k < 2^28: every index and the SHA-256 of the
    (k, packed key) records;
  - the first 3 symbols over 2^24 candidates from k = 2^32 + 12345, reached by
    the third launch: indices past 32 bits;
  - a 5-symbol suffix and a 5-symbol substring over k < 2^28 (the address
    kernel, the sign of x formed in the loop and derived there);
  - the dictionary workload (10,000 prefixes of 5-8 symbols behind a 24-bit
    bitmap and the exact table) over k < 2^28.
Each case prints itself before it touches the device.
"""
import hashlib
import json
import os

import pytest

from onion_gen_amd import filters, gpu, key, workload

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FULL = dict(threads=1 << 20, half=1024, rounds=1, hit_cap=1 << 20)


def say(msg):
    print(msg, flush=True)


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def window(name):
    return next(w for w in golden("config2_scale.json")["windows"] if w["name"] == name)


def hit_digest(hits):
    h = hashlib.sha256()
    for k, packed in sorted(hits):
        h.update(k.to_bytes(8, "little") + packed)
    return h.hexdigest()


def engine_window(seed, fs, k0, n):
    """Survivors of the default-shape engine with k0 <= k < k0 + n, sorted
    (k, packed), from as many launches as reach k0 + n."""
    eng = gpu.Engine.with_filters(key.start_point(seed), fs, **FULL)
    got, done, launches = [], 0, 0
    try:
        assert eng.shape() == (FULL["threads"], FULL["half"], FULL["rounds"])
        while done < k0 + n:
            for s in eng.advance():
                assert s.offset % 8 == 0
                k = s.offset // 8
                if k0 <= k < k0 + n:
                    got.append((k, s.packed))
            done += eng.batch()
            launches += 1
    finally:
        eng.close()
    say(f"  {launches} launch(es), {len(got)} survivors in the window")
    return sorted(got)


def seed2():
    return bytes.fromhex(golden("config2_scale.json")["seed"])


def check_full(w):
    seed = seed2()
    say(f"case: '{w['filter']}' over [{w['k0']}, +{w['n']}) at {FULL}, every hit")
    got = engine_window(seed, filters.FilterSet.parse(w["filter"]), w["k0"], w["n"])
    want = [(h["k"], bytes.fromhex(h["pk"])) for h in w["hits"]]
    assert [k for k, _ in got] == [k for k, _ in want]
    for (k, packed), (_, pk) in zip(got, want):
        assert packed == pk[:31] + bytes([pk[31] & 0x7F]), k
    assert hit_digest(got) == w["sha256"]
    # the hit sink re-derives each public key, sign included, from its secret
    recs = key.finish_hits(seed, [gpu.Survivor(8 * k, p) for k, p in got])
    for rec, (k, pk), h in zip(recs, want, w["hits"]):
        assert rec.status == key.HIT_OK and rec.public_key == pk, k
        assert rec.hostname == h["address"] + ".onion"


def check_indices(w, seed=None, fs=None):
    seed = seed or seed2()
    fs = fs or filters.FilterSet.parse(w["filter"])
    say(f"case: '{w.get('filter', 'dictionary')}' over [{w['k0']}, +{w['n']}) at {FULL}, "
        f"{w['count']} hits by index and digest")
    got = engine_window(seed, fs, w["k0"], w["n"])
    ks = [k for k, _ in got]
    if ks != w["ks"]:
        missing = sorted(set(w["ks"]) - set(ks))[:10]
        extra = sorted(set(ks) - set(w["ks"]))[:10]
        raise AssertionError(f"{len(ks)} hits, want {len(w['ks'])}: missing {missing}, "
                             f"extra {extra}")
    assert len(got) == w["count"]
    assert hit_digest(got) == w["sha256"]


def test_prefix4_over_2p28_every_hit():
    check_full(window("prefix4"))


def test_prefix3_over_2p28_by_digest():
    check_indices(window("prefix3"))


def test_prefix3_past_2p32():
    w = window("prefix3_past_2p32")
    assert w["k0"] > 1 << 32
    check_full(w)


def test_suffix5_over_2p28_by_digest():
    check_indices(window("suffix5"))


def test_contains5_over_2p28_by_digest():
    check_indices(window("contains5"))


def test_dictionary_over_2p28_by_digest():
    g = golden("config3_dictionary.json")
    assert g["dictionary"] == "workload.dictionary(10000, 2)"
    fs = filters.FilterSet.parse(workload.dictionary(10000, 2), index_bits=g["index_bits"])
    assert fs.index_words()[1] == 24
    check_indices(g, seed=bytes.fromhex(g["seed"]), fs=fs)
