"""Writes the golden fixtures from the CPU oracle (run from the repo root):

    python tests/golden/make_golden.py

config1_hits.json: every hit of the CPU-sized workload - the secret seed is
ChaCha20 seeded 0 (32 bytes), the filter a 4-symbol prefix drawn from ChaCha20
seeded 1, BatchEngine width 256, candidates k = 0 .. 2^24-1.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from onion_gen_amd import workload  # noqa: E402
from tests.oracle_lib import OracleFilters, digest, search, start_point  # noqa: E402


def main():
    c1 = workload.config1()
    start = start_point(c1["seed"])
    t = time.time()
    hits = search(start, 0, c1["candidates"], c1["width"], OracleFilters(c1["prefix"]))
    x, s = digest(start, 0, 1 << 20, 256)
    out = {
        "seed": c1["seed"].hex(), "prefix": c1["prefix"], "width": c1["width"],
        "candidates": c1["candidates"],
        "hits": [{"k": k, "pk": pk.hex(), "address": a} for k, pk, a in hits],
        "digest_2p20": {"xor": x, "sum": s},
    }
    with open(os.path.join(ROOT, "tests", "golden", "config1_hits.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(hits)} hits in {time.time() - t:.1f}s")


if __name__ == "__main__":
    main()
