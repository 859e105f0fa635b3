"""The kernel-level radix bench: which field representation the search uses.

Runs og_hip_radix_bench for each representation, checks every chain end
against Python big-integer arithmetic, prints field multiplications per
second and records the fastest in the tuning cache under
``device/<name>/hip-radix/0/<version>`` (src/tuning.rs:65-74's key shape,
field ``radix``).

    python -m onion_gen_amd.radix_bench [--threads N] [--iters N]
"""
import argparse
import ctypes
import json

from ._lib import TuningShape, err_buf, lib
from . import gpu, tuning

P = 2 ** 255 - 19
RADIXES = {5: "5 x 51 in 64-bit limbs (fiat-crypto form)", 8: "8 x 32 saturated, row-wise",
           9: "9 x 29, carry-free columns", 10: "10 x 25.5 (ref10 shape)"}
ISSUE_OPS = {0: "v_mad_u64_u32", 1: "v_mul_lo_u32", 2: "v_lshrrev_b64", 3: "v_add_u32",
             4: "v_bitop3_b32", 5: "v_alignbit_b32",
             6: "v_mad_u64_u32, one dependent chain, 3 waves per SIMD"}


def radix_input(lane, which, limb):
    m = 0xFFFFFFFF
    x = (lane * 0x9E3779B9 + which * 0x85EBCA6B + limb * 0xC2B2AE35 + 0x27220A95) & m
    x ^= x >> 15
    x = (x * 0x2C1B3C6D) & m
    x ^= x >> 12
    x = (x * 0x297A2D39) & m
    x ^= x >> 15
    return x & 0x7FFFFFFF if limb == 7 else x


def lane_value(lane, which):
    return sum(radix_input(lane, which, l) << (32 * l) for l in range(8))


def run(radix, threads, iters, check_lanes=64):
    rate = ctypes.c_double()
    out = ctypes.create_string_buffer(threads * 64)
    e = err_buf()
    rc = lib().og_hip_radix_bench(0, radix, threads, iters, ctypes.byref(rate), out, e, len(e))
    if rc != 0:
        raise RuntimeError(e.value.decode(errors="replace"))
    for lane in range(0, threads, max(1, threads // check_lanes)):
        y = pow(lane_value(lane, 2), iters, P)
        for which in (0, 1):
            want = lane_value(lane, which) * y % P
            got = int.from_bytes(out.raw[lane * 64 + 32 * which:lane * 64 + 32 * which + 32], "little")
            if got != want:
                raise AssertionError(f"radix {radix} lane {lane} chain {which}: wrong result")
    return rate.value


def issue_rate(op=0, threads=1 << 20, iters=2048, device=0):
    """og_hip_issue_bench: (lane-instructions per second, shader clock in Hz)
    of one VALU instruction kind (ISSUE_OPS) issued back to back."""
    rate, clock = ctypes.c_double(), ctypes.c_double()
    e = err_buf()
    if lib().og_hip_issue_bench(device, op, threads, iters, ctypes.byref(rate),
                                ctypes.byref(clock), e, len(e)) != 0:
        raise RuntimeError(e.value.decode(errors="replace"))
    return rate.value, clock.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=1 << 18)
    ap.add_argument("--iters", type=int, default=256)
    ap.add_argument("--no-record", action="store_true")
    ap.add_argument("--issue", action="store_true",
                    help="also the issue rate of each VALU instruction the product is made of")
    args = ap.parse_args()
    if args.issue:
        for op, name in ISSUE_OPS.items():
            best = max((issue_rate(op) for _ in range(3)), key=lambda rc: rc[0])
            print(json.dumps({"instruction": name, "lane_ops_per_s": best[0],
                              "clock_hz": best[1],
                              "per_lane_per_clock": best[0] / (256 * 64 * best[1])
                              if best[1] else None}), flush=True)
    results = {}
    for radix, name in RADIXES.items():
        rate = run(radix, args.threads, args.iters)
        results[radix] = rate
        print(json.dumps({"radix": radix, "form": name, "field_mul_per_s": rate,
                          "checked": True}), flush=True)
    best = max(results, key=results.get)
    print(json.dumps({"best": best, "form": RADIXES[best], "field_mul_per_s": results[best]}))
    if not args.no_record:
        dev = gpu.look()[0]
        key = tuning.device_key(dev.name, "hip-radix", 0).encode()
        lib().og_tuning_put(key, ctypes.byref(TuningShape(0, 0, results[best], 0, 0, best)))


if __name__ == "__main__":
    main()
