"""The serve loop on the device: requests arrive, queue, run in slots and
complete at their n-th hit; every reported hit is a real one and a request's
hits over the candidates it examined are exactly the oracle's."""
import hashlib
import re

import pytest

from onion_gen_amd import batch, filters, key, serve, workload
from tests.oracle_lib import OracleFilters, search

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def clear_sign(pk):
    return pk[:31] + bytes([pk[31] & 0x7F])


def regex_of(text):
    if text.startswith("suffix:"):
        return re.compile(text[7:].replace("?", ".") + "$")
    if text.startswith("contains:"):
        return re.compile(text[9:].replace("?", "."))
    return re.compile("^" + text.replace("?", "."))


def test_requests_complete_with_real_hits():
    say("case: serve 12 easy requests on 4 slots x 512 chains, half 64")
    texts = ["2x", "suffix:aqd", "contains:a2b", "q7", "suffix:yid", "contains:zz"] * 2
    wants = [3, 5, 2, 4, 3, 2] * 2
    reqs = []
    for i, (t, w) in enumerate(zip(texts, wants)):
        reqs.append((0.05 * i, hashlib.sha256(b"req%d" % i).digest(), t, w))
    st, res = serve.run(reqs, duration=60.0, slots=4, chains=512, half=64, max_queue=64)
    say(f"stats: {st}")
    assert st["completed"] == len(reqs), st
    assert st["p99_time_to_n"] >= st["p50_time_to_n"] > 0
    for (t_arr, seed, text, want), r in zip(reqs, res):
        assert r["status"] == "done" and r["hits"] >= want
        for offset, packed in r["found"]:
            pk, addr = key.finish_hit(seed, offset)
            assert clear_sign(pk) == packed
            assert regex_of(text).search(addr), (text, addr)
        # every hit over what the request examined was found; it recorded
        # exactly the number it wanted, each one real
        oracle_hits = search(key.start_point(seed), 0, r["examined"], 256, OracleFilters(text))
        want_hits = [(k, clear_sign(pk)) for k, pk, _ in oracle_hits]
        assert r["seen"] == len(want_hits) and r["hits"] == want
        assert set((o // 8, p) for o, p in r["found"]) <= set(want_hits)


def test_bounded_queue_rejects_overflow():
    """A device budget that holds one search: the first arrival runs, two wait
    in the queue of 2, the other seven are turned away."""
    say("case: serve 10 simultaneous arrivals, budget for one search, queue of 2")
    reqs = [(0.0, hashlib.sha256(b"q%d" % i).digest(), "2", 1) for i in range(10)]
    with batch.Batch(slots=1, chains=256, half=32) as b:
        one = b.footprint(filters.FilterSet.parse("2"))
    assert one > 0
    st, res = serve.run(reqs, duration=30.0, slots=1, chains=256, half=32, max_queue=2,
                        device_budget=one)
    assert st["rejected"] == 7 and st["completed"] == 3, st
    assert st["max_resident"] == 1, st
    assert [r["status"] for r in res].count("rejected") == 7


def test_resident_searches_outnumber_slots():
    """With room on the device every arrival becomes resident at once and the
    slots are time-shared: ten searches on two slots all finish."""
    say("case: serve 10 simultaneous arrivals, 2 slots, ample budget")
    reqs = [(0.0, hashlib.sha256(b"r%d" % i).digest(), "2xe", 2) for i in range(10)]
    # a queue of 0: nothing may wait, so every arrival must fit at once
    st, res = serve.run(reqs, duration=60.0, slots=2, chains=256, half=32, max_queue=0)
    assert st["completed"] == 10 and st["rejected"] == 0, st
    assert st["max_resident"] == 10, st
    for (_, seed, text, want), r in zip(reqs, res):
        assert r["status"] == "done" and r["hits"] >= want
        got = sorted((o // 8, p) for o, p in r["found"])
        oracle_hits = search(key.start_point(seed), 0, r["examined"], 256, OracleFilters(text))
        want_hits = [(k, clear_sign(pk)) for k, pk, _ in oracle_hits]
        assert r["seen"] == len(want_hits) and r["hits"] == want
        assert set(got) <= set(want_hits)


def test_mix_generators_are_reachable_filters():
    mix = workload.search_mix(64, 4)
    kinds = {t.split(":")[0] if ":" in t else "prefix" for _, t, _ in mix}
    assert kinds == {"prefix", "suffix", "contains"}
    assert all(1 <= w <= 1000 for _, _, w in mix)
