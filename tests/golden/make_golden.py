"""Writes the golden fixtures from the CPU oracle (run from the repo root):

    python tests/golden/make_golden.py [config1] [config2] [config3] [mix64]

config1_hits.json: every hit of the CPU-sized workload - the secret seed is
ChaCha20 seeded 0 (32 bytes), the filter a 4-symbol prefix drawn from ChaCha20
seeded 1, BatchEngine width 256, candidates k = 0 .. 2^24-1.

config2_scale.json: the single-GPU workload's seed (the same ChaCha20 seeded
0) over windows the device's default launch covers, so that the full-size
launch is checked against the oracle itself:
  - the 6-symbol prefix's first 4 symbols over k < 2^28 (about 256 hits), every
    hit with its public key and address;
  - its first 3 symbols over k < 2^28 (about 8,192 hits): every index, and a
    SHA-256 over the (k, packed key) records;
  - its first 3 symbols over 2^24 candidates from k = 2^32 + 12345 (past 32 bits
    of candidate index), every hit;
  - a 5-symbol suffix and a 5-symbol substring over k < 2^28 (the checksum
    forms: the address of every candidate decides them), every index and the
    SHA-256.

config3_dictionary.json: the dictionary workload - 10,000 distinct prefixes of
5-8 symbols (ChaCha20 seeded 2) behind a 24-bit bitmap - over k < 2^28 from the
same seed: every index and the SHA-256.

mix64_golden.json: the batched workload - the 64 searches of the mix (seed 4:
own seed, prefix 4-7 / suffix 3-5 / substring 5, wanted hits log-uniform in
1..1000) - each over the candidates two steps of the slot engine's default
unit (16384 chains x 513 candidates each) examine: the count of hits and the
SHA-256 over the (k, packed key) records.

The SHA-256 of a hit list is over its records in increasing k, each the index
as 8 little-endian bytes then the packed key with the sign bit cleared (the
form the device reports).
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from onion_gen_amd import workload  # noqa: E402
from tests.oracle_lib import OracleFilters, digest, search, search_mt, start_point  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")

# The slot engine's default unit: 16384 chains of 2 * 256 + 1 candidates.
MIX_PER_STEP = 16384 * 513
MIX_STEPS = 2


def hit_digest(hits):
    """SHA-256 over (k as 8 LE bytes, packed key with the sign bit clear)."""
    h = hashlib.sha256()
    for k, pk in sorted((k, pk) for k, pk, *_ in hits):
        h.update(k.to_bytes(8, "little") + pk[:31] + bytes([pk[31] & 0x7F]))
    return h.hexdigest()


def _write(name, obj):
    with open(os.path.join(GOLDEN, name), "w") as f:
        json.dump(obj, f, indent=1)


def config1():
    c1 = workload.config1()
    start = start_point(c1["seed"])
    t = time.time()
    hits = search(start, 0, c1["candidates"], c1["width"], OracleFilters(c1["prefix"]))
    x, s = digest(start, 0, 1 << 20, 256)
    _write("config1_hits.json", {
        "seed": c1["seed"].hex(), "prefix": c1["prefix"], "width": c1["width"],
        "candidates": c1["candidates"],
        "hits": [{"k": k, "pk": pk.hex(), "address": a} for k, pk, a in hits],
        "digest_2p20": {"xor": x, "sum": s},
    })
    print(f"config1: {len(hits)} hits in {time.time() - t:.1f}s", flush=True)


def _window(start, name, text, k0, n, full):
    t = time.time()
    hits = search_mt(start, k0, n, 256, OracleFilters(text), cap=1 << 20)
    out = {"name": name, "filter": text, "k0": k0, "n": n, "count": len(hits),
           "sha256": hit_digest(hits)}
    if full:
        out["hits"] = [{"k": k, "pk": pk.hex(), "address": a} for k, pk, a in hits]
    else:
        out["ks"] = [k for k, _, _ in hits]
    print(f"config2 {name} '{text}' [{k0}, +{n}): {len(hits)} hits in {time.time() - t:.1f}s",
          flush=True)
    return out


def config2():
    c2 = workload.config2()
    start = start_point(c2["seed"])
    p = c2["prefix"]
    windows = [
        _window(start, "prefix4", p[:4], 0, 1 << 28, True),
        _window(start, "prefix3", p[:3], 0, 1 << 28, False),
        _window(start, "prefix3_past_2p32", p[:3], (1 << 32) + 12345, 1 << 24, True),
        # 5-symbol suffix: the last two symbols are the version's (d) and one
        # of the four that precede it, so 1 key in 2^15 * 4 matches
        _window(start, "suffix5", "suffix:" + p[1:4] + "ad", 0, 1 << 28, False),
        # 5-symbol substring at any of its 52 placements
        _window(start, "contains5", "contains:" + p[2:6] + "q", 0, 1 << 28, False),
    ]
    _write("config2_scale.json", {"seed": c2["seed"].hex(), "prefix": p, "windows": windows})


def config3():
    seed = workload.chacha_stream(0, 32)
    start = start_point(seed)
    words = workload.dictionary(10000, 2)
    t = time.time()
    hits = search_mt(start, 0, 1 << 28, 256, OracleFilters(words), cap=1 << 20)
    _write("config3_dictionary.json", {
        "seed": seed.hex(), "dictionary": "workload.dictionary(10000, 2)", "index_bits": 24,
        "k0": 0, "n": 1 << 28, "count": len(hits), "sha256": hit_digest(hits),
        "ks": [k for k, _, _ in hits]})
    print(f"config3: {len(hits)} hits in {time.time() - t:.1f}s", flush=True)


def mix64():
    n = MIX_STEPS * MIX_PER_STEP
    out = []
    t = time.time()
    for i, (seed, text, want) in enumerate(workload.search_mix(64, 4)):
        hits = search_mt(start_point(seed), 0, n, 256, OracleFilters(text), cap=1 << 20)
        out.append({"search": i, "filter": text, "want": want, "count": len(hits),
                    "sha256": hit_digest(hits)})
        print(f"mix64 {i} '{text}': {len(hits)} hits ({time.time() - t:.0f}s)", flush=True)
    _write("mix64_golden.json", {"mix": "workload.search_mix(64, 4)", "steps": MIX_STEPS,
                                 "per_step": MIX_PER_STEP, "n": n, "searches": out})


def main():
    which = sys.argv[1:] or ["config1", "config2", "config3", "mix64"]
    for name in which:
        globals()[name]()


if __name__ == "__main__":
    main()
