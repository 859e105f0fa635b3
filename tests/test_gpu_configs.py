"""The batched and served workloads under test at their stated shapes.

Batched: the 64 searches of the mix (workload.search_mix(64, 4): each its own
seed; prefix 4-7, suffix 3-5 or substring 5; wanted hits log-uniform in
1..1000) resident together on the slot engine's default unit (64 slots x 16384
chains x 256 pairs per window), two steps, so that each search examines
k < 2 x 16384 x 513.  The oracle's hits over that range are in
tests/golden/mix64_golden.json (count and SHA-256 per search):
  - with each search's quota set to its wanted hits, every search counts
    exactly the oracle's hits, records min(want, count) of them, and every
    recorded hit is one of the oracle's (checked candidate by candidate);
  - with no quota, every search records exactly the oracle's hit set.

Served: a prefix of the served workload's arrivals (Poisson 200 requests/s,
workload.poisson_arrivals(200, 120, 5); filters of the mix with n in
{1, 10, 100}, workload.search_mix(., 6, hits="choice")) through og_serve_run.
Which n of a request's hits it records depends on the order in which the
waves of its last step report them, so the result that must be the same alone
and in the mix is the hit set: for each request, the hits it counted equal the
single-search engine's over the candidates it examined (a prefix of its
chain), every hit it recorded is one of the oracle's, a finished request
recorded exactly n, and run alone it finishes with n of the same hits.
"""
import json
import os

import pytest

from onion_gen_amd import batch, filters, gpu, key, serve, workload
from tests.oracle_lib import OracleFilters, search

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)


def mix_golden():
    with open(os.path.join(GOLDEN, "mix64_golden.json")) as f:
        return json.load(f)


class OracleCheck:
    """Candidate-by-candidate oracle: is k a hit of the search, and is its
    packed key (sign clear) this one."""

    def __init__(self, seed, text):
        self.start = key.start_point(seed)
        self.fs = OracleFilters(text)

    def __call__(self, k, packed):
        got = search(self.start, k, 1, 1, self.fs)
        if len(got) != 1:
            return False
        pk = got[0][1]
        return got[0][0] == k and packed == pk[:31] + bytes([pk[31] & 0x7F])


def run_mix(with_quota, ring_cap):
    import hashlib
    mix = workload.search_mix(64, 4)
    g = mix_golden()
    assert g["mix"] == "workload.search_mix(64, 4)" and len(g["searches"]) == 64
    recorded = [[] for _ in mix]
    with batch.Batch(slots=64, chains=16384, half=256, rounds=1, ring_cap=ring_cap) as b:
        assert b.per_step() == g["per_step"]
        ids = []
        for i, (seed, text, want) in enumerate(mix):
            sid = b.admit(key.start_point(seed), filters.FilterSet.parse(text), i)
            if with_quota:
                b.set_want(sid, want)
            ids.append(sid)
        for step in range(g["steps"]):
            hits = b.step()
            say(f"  step {step}: {len(hits)} records")
            for h in hits:
                recorded[h.tag].append((h.offset // 8, h.packed))
        states = [b.search_state(sid) for sid in ids]
        found = [b.found(sid) for sid in ids]
    for i, (st, gs) in enumerate(zip(states, g["searches"])):
        assert gs["filter"] == mix[i][1] and gs["want"] == mix[i][2]
        assert st["examined"] == g["n"], (i, st)
        assert found[i] == gs["count"], (i, gs["filter"], found[i], gs["count"])
        assert all(k < g["n"] for k, _ in recorded[i])
    return mix, g, recorded, hashlib


def test_config4_mix64_with_quotas():
    say("case: 64-search mix on 64 x 16384 x 256, quotas at the wanted hits, 2 steps")
    mix, g, recorded, _ = run_mix(True, 1 << 18)
    checked = 0
    for i, ((seed, text, want), gs) in enumerate(zip(mix, g["searches"])):
        got = recorded[i]
        assert len(got) == min(want, gs["count"]), (i, text, len(got), want, gs["count"])
        assert len(set(got)) == len(got)
        check = OracleCheck(seed, text)
        for k, packed in got:
            assert check(k, packed), (i, text, k)
            checked += 1
    say(f"  {checked} recorded hits each confirmed by the oracle")


def test_config4_mix64_every_hit():
    say("case: 64-search mix on 64 x 16384 x 256, no quota, 2 steps")
    mix, g, recorded, hashlib = run_mix(False, 1 << 21)
    for i, gs in enumerate(g["searches"]):
        h = hashlib.sha256()
        for k, packed in sorted(recorded[i]):
            h.update(k.to_bytes(8, "little") + packed)
        assert len(recorded[i]) == gs["count"], (i, gs["filter"])
        assert h.hexdigest() == gs["sha256"], (i, gs["filter"])


def served_requests(window_s):
    arrivals = workload.poisson_arrivals(200, 120, 5)
    mix = workload.search_mix(len(arrivals), 6, hits="choice")
    return [(t, seed, text, want) for t, (seed, text, want) in zip(arrivals, mix)
            if t < window_s]


def engine_count(seed, text, n):
    """Hits of the single-search engine over k < n."""
    eng = gpu.Engine.with_filters(key.start_point(seed), filters.FilterSet.parse(text),
                                  threads=1 << 16, half=256, rounds=1, hit_cap=1 << 20)
    count, done = 0, 0
    try:
        while done < n:
            count += sum(1 for s in eng.advance() if s.offset // 8 < n)
            done += eng.batch()
    finally:
        eng.close()
    return count


def check_request(req, res, label):
    t, seed, text, want = req
    assert res["status"] in ("done", "running"), (label, res["status"])
    n = min(res["hits"], 16)
    assert len(res["found"]) == n
    check = OracleCheck(seed, text)
    for offset, packed in res["found"]:
        assert offset % 8 == 0 and offset // 8 < res["examined"], (label, offset)
        assert check(offset // 8, packed), (label, text, offset)
    if res["status"] == "done":
        assert res["hits"] == want, (label, res["hits"], want)
        assert res["seen"] >= want
    else:
        assert res["hits"] == res["seen"] < want
    assert res["seen"] == engine_count(seed, text, res["examined"]), (label, text)


def test_config5_served_poisson_prefix():
    reqs = served_requests(1.0)
    say(f"case: serve loop, the first {len(reqs)} Poisson arrivals (1 s of 200 req/s), 4 s, "
        "64 x 16384 x 256")
    st, res = serve.run(reqs, duration=4.0, slots=64, chains=16384, half=256,
                        max_queue=1 << 16, policy=batch.LEAST_SERVED)
    say(f"  completed {st['completed']}, running {st['running']}, queued {st['queued']}, "
        f"rejected {st['rejected']}, p99 time-to-n {st['p99_time_to_n']:.3f} s")
    assert st["rejected"] == 0 and st["queued"] == 0
    assert st["completed"] >= len(reqs) // 3
    for i, (req, r) in enumerate(zip(reqs, res)):
        check_request(req, r, f"request {i}")
    # alone: a few of the finished requests, each the only one on the device
    done = [i for i, r in enumerate(res) if r["status"] == "done"]
    kinds = {}
    for i in done:
        kinds.setdefault(reqs[i][2].split(":")[0] if ":" in reqs[i][2] else "prefix", i)
    for kind, i in sorted(kinds.items()):
        t, seed, text, want = reqs[i]
        say(f"case: request {i} '{text}' (n = {want}) alone")
        st1, res1 = serve.run([(0.0, seed, text, want)], duration=20.0, slots=64, chains=16384,
                              half=256, max_queue=4, policy=batch.LEAST_SERVED)
        assert res1[0]["status"] == "done"
        check_request((0.0, seed, text, want), res1[0], f"request {i} alone")


def test_serve_rejects_what_never_fits_and_serves_the_rest():
    """Under a budget of exactly one literal prefix's footprint: a suffix
    request (its automaton tables make it larger) can never fit and is
    rejected on arrival, and the prefix requests behind it are served one after
    the other; a request that asks for no hit, or whose filter text fills its
    64 bytes with no terminator, is rejected too.  None holds up the queue."""
    seed = workload.chacha_stream(9, 32)
    with batch.Batch(slots=4, chains=2048, half=64) as b:
        small = b.footprint(filters.FilterSet.parse("2x"))
        big = b.footprint(filters.FilterSet.parse("suffix:aqd"))
    say(f"case: serve loop, budget {small} bytes (a suffix request takes {big})")
    assert big > small > 0
    reqs = [(0.0, seed, "2x", 2),
            (0.0, seed, "suffix:aqd", 1),
            (0.01, seed, "2y", 0),
            (0.02, workload.chacha_stream(10, 32), "2z", 2),
            (0.03, seed, "a" * 64, 1)]
    st, res = serve.run(reqs, duration=20.0, slots=4, chains=2048, half=64, max_queue=8,
                        device_budget=small)
    assert [r["status"] for r in res] == ["done", "rejected", "rejected", "done", "rejected"]
    assert st["completed"] == 2 and st["rejected"] == 3 and st["queued"] == 0
    assert st["wall_s"] < 19.0  # nothing left waiting: the loop ends with the work
    for req, r in zip(reqs, res):
        if r["status"] == "done":
            check_request(req, r, req[2])
