"""Candidate-key search throughput on one MI355X.

Workload (the single-GPU configuration of the headline metric): one 6-symbol
base32 prefix drawn from ChaCha20 seeded 1, the secret seed from ChaCha20
seeded 0, 2^34 sequential candidates P0 + k*8G.  A step is one engine advance
(one launch over threads * (2*half+1) * rounds candidates); the default step
count covers 2^34 candidates.  Every survivor is re-derived from its secret
scalar on the host and checked against the prefix.

Prints one JSON line (rank 0).  `value` is candidates/s with the start state
already on the device; `roofline` prices the search kernel's field
multiplications against the VALU 32x32->64 multiply-add issue rate;
`cpu_baseline` is the oracle's BatchEngine (the reference's CPU engine restated
in C) on the host cores, on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs x 4 SIMD x 16 lanes at 2.4 GHz = one VALU op per lane per clock.
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9


def field_muls_per_candidate(half, with_x=False):
    """Field multiplications (squarings included) the kernel performs per
    candidate: 3 per pair forward (C = T dq and two prefix products), 8 per
    pair backward (two inverses peeled, two products A and B, two y), the
    centre (4), one inversion (254 sq + 11 mul) and the stride addition (7)
    per window of 2*half+1 candidates; +2 per candidate when the sign of x is
    needed.  Priced at 64 32x32-bit products each, the radix-independent count
    of an 8 x 32 schoolbook."""
    window = 11 * half + 4 + 265 + 7
    return window / (2 * half + 1) + (2 if with_x else 0)


def cpu_baseline(prefix, seconds):
    from tests.oracle_lib import OracleFilters, load
    import ctypes
    from onion_gen_amd import workload
    L = load()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    fs = OracleFilters(prefix)
    examined = ctypes.c_uint64()
    rate = L.obench_batch(workload.chacha_stream(0, 32), 256, cores, seconds, fs.h,
                          ctypes.byref(examined))
    return {"value": rate, "unit": "keys/s", "cores": cores, "kind": "port",
            "sample": f"oracle BatchEngine width 256 (src/batch.rs restated in C), one chain per "
                      f"thread, prefix '{prefix}', {seconds:.0f} s wall, {examined.value} candidates"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="0: cover 2^34 candidates")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--threads", type=int, default=1 << 18)
    ap.add_argument("--half", type=int, default=256)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a 1 GiB device-to-device copy: the known-size kernel the "
                         "FETCH_SIZE / WRITE_SIZE counters are scaled by")
    args = ap.parse_args()
    if args.copy_calibrate:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the search has no cross-GPU step; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from onion_gen_amd import filters, gpu, key, workload

    c2 = workload.config2()
    prefix, seed = c2["prefix"], c2["seed"]
    start = key.start_point(seed)
    fs = filters.FilterSet.parse(prefix)
    eng = gpu.Engine.with_filters(start, fs, threads=args.threads, half=args.half,
                                  rounds=args.rounds)
    batch = eng.batch()
    steps = args.steps or -(-c2["candidates"] // batch)
    survivors = []
    for _ in range(args.warmup):
        survivors += eng.advance()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(steps):
        survivors += eng.advance()
        kernel_ms += eng.last_ms()
    wall = time.perf_counter() - t0
    examined = steps * batch
    rate = examined / wall

    verified = 0
    for s in survivors:
        got = key.finish_hit(seed, s.offset)
        if got is not None and got[1].startswith(prefix) and \
                got[0][:31] == s.packed[:31] and (got[0][31] & 0x7F) == s.packed[31]:
            verified += 1

    F = field_muls_per_candidate(args.half)
    kernel_rate = examined / (kernel_ms * 1e-3)
    achieved = kernel_rate * F * 64  # 32x32->64 partial products per second
    out = {
        "metric": "candidate keys/sec on 1 MI355X (bit-exact hits)",
        "value": rate,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"one 6-symbol prefix '{prefix}', {examined} sequential "
                               f"candidates (2^34 at the default step count)",
                   "threads": args.threads, "half": args.half, "rounds": args.rounds,
                   "candidates_per_step": batch},
        "hits": len(survivors),
        "hits_rederived": verified,
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": VALU_PEAK_LANE_OPS / 1e12,
            "unit": "T mad32/s",
            "frac": achieved / VALU_PEAK_LANE_OPS,
            "traffic": None,
            "kernel_avg_ms": kernel_ms / steps,
            "field_muls_per_candidate": F,
        },
    }
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prefix, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
