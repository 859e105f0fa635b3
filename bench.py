"""Candidate-key search throughput on one MI355X.

Workload (the single-GPU configuration of the headline metric): one 6-symbol
base32 prefix drawn from ChaCha20 seeded 1, the secret seed from ChaCha20
seeded 0, 2^34 sequential candidates P0 + k*8G.  A step is one engine advance
(one launch over threads * (2*half+1) * rounds candidates); the default step
count covers 2^34 candidates.  Every survivor goes through the hit sink (its
secret at the offset, its public key re-derived from the scalar alone) and its
address is matched on the host; the run exits non-zero if one fails.  (With
--filter, a profiling run of another form, at most 65,536 evenly spaced
survivors are checked, and the line says how many.)

Prints one JSON line (rank 0).  `value` is candidates/s with the start state
already on the device; `roofline` prices the search kernel's field
multiplications against the VALU 32x32->64 multiply-add issue rate;
`cpu_baseline` is the oracle's PairEngine (the reference's default CPU engine
restated in C) on the host cores, on a bounded sample, with its BatchEngine
beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs x 4 SIMD x 16 lanes at the published 2.4 GHz peak clock, one
# VALU op per lane per clock: the nominal ceiling, reported beside the
# measured peak that prices the roofline.
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9


def progress(msg):
    """A line on stderr as each part of the run ends (the JSON line stays alone
    on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def measured_mad_peak():
    """P_mul: the peak v_mad_u64_u32 issue rate measured on this device now
    (og_hip_issue_bench, eight independent chains per lane over 2^20 lanes),
    the best of three, with the shader clock it ran at."""
    from onion_gen_amd.radix_bench import issue_rate
    return max((issue_rate(0, 1 << 20, 2048) for _ in range(3)), key=lambda rc: rc[0])


def field_muls_per_candidate(half, key_form=True, with_x=False):
    """Field multiplications (squarings included) the kernel performs per
    candidate: 2 per pair forward (the shared denominator E = ZZ + TT(-dq^2)
    and the prefix product), 8 per pair backward (two inverses peeled, the
    double products alpha and beta, two each, and the two y), and per window
    of 2*half+1 candidates the six products of the centre's coordinates, the
    centre (4), one inversion and the stride addition (7).  The inversion is
    the divstep recurrence, whose 20 batches make 90 products of 30-bit limbs
    each (the updates of f, g, d and e): 1,800 products, 28.1 multiplications'
    worth; its divsteps are logic, not products.
    The key forms make only three columns of each y (30 of its 100 products,
    the probe word of fe10_mul_low26), so their pair counts 8.6; the whole y
    of the rare lanes whose probe passes is left out.  +3 per candidate when
    the sign of x is needed (gamma and delta, two each, and the two x per
    pair).  Priced at 64 32x32-bit products each, the radix-independent count
    of an 8 x 32 schoolbook."""
    pair = 8.6 if key_form else 10.0
    window = pair * half + 6 + 4 + 1800 / 64 + 7
    return window / (2 * half + 1) + (3 if with_x else 0)


def measured_traffic(threads, half, rounds):
    """HBM bytes per launch of the search kernel at this shape, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy
    calibration it records), or None when that summary is for another shape."""
    path = os.path.join(ROOT, "profiles", "search_prefix6_traffic.md")  # lists each shape measured
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    want = f"traffic: threads={threads} half={half} rounds={rounds} bytes_per_launch="
    for line in lines:
        if line.startswith(want):
            return int(line[len(want):])
    return None


def _pct(xs, q):
    if not xs:
        return None
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(q * len(xs)))]


def run_mix(searches, seconds, slots, chains, half, policy, ring_cap=1 << 20):
    """The searches (seed32, filter text, wanted hits) resident on one slot
    engine for `seconds` of wall time.  A search retires at its target.
    Returns aggregate candidates/s and per-search first-hit / done times."""
    from onion_gen_amd import batch, filters, key
    import numpy as np
    b = batch.Batch(slots=slots, chains=chains, half=half, rounds=1, ring_cap=ring_cap,
                    policy=policy)
    tag_of = {}
    for i, (seed, text, want) in enumerate(searches):
        sid = b.admit(key.start_point(seed), filters.FilterSet.parse(text), i)
        b.set_want(sid, want)
        tag_of[sid] = i
    hits = [0] * len(searches)
    first, done = [None] * len(searches), [None] * len(searches)
    live = set(tag_of)
    steps = 0
    device_ms = 0.0
    t0 = time.perf_counter()
    while live and time.perf_counter() - t0 < seconds:
        tags = b.step_tags()
        now = time.perf_counter() - t0
        steps += 1
        device_ms += b.last_ms()
        if len(tags):
            for t, c in zip(*np.unique(tags, return_counts=True)):
                hits[int(t)] += int(c)
                if first[int(t)] is None:
                    first[int(t)] = now
        for sid in [s for s in live if hits[tag_of[s]] >= searches[tag_of[s]][2]]:
            done[tag_of[sid]] = now
            b.retire(sid)
            live.discard(sid)
    wall = time.perf_counter() - t0
    examined = b.examined()
    b.close()
    got = [t for t in first if t is not None]
    return {"keys_per_s": examined / wall, "device_keys_per_s": examined / (device_ms * 1e-3),
            "window_s": wall, "steps": steps, "completed": sum(d is not None for d in done),
            "with_hit": len(got), "p50_first_hit_s": _pct(got, 0.5),
            "p99_first_hit_s": _pct(got, 0.99)}


def batched_leg(seconds):
    """The 64-search mix in one launch per step, against one search per
    launch: at the same per-search shape, and at the full-GPU shape."""
    from onion_gen_amd import batch, filters, workload
    mix = workload.search_mix(64, 4)
    kinds = {}
    for _, text, _ in mix:
        k = text.split(":")[0] if ":" in text else "prefix"
        kinds[k] = kinds.get(k, 0) + 1
    # the unit shape the slot engine's timed walk (og_batch_tune) picks for
    # this mix: slots x chains per slot x pairs per window
    t0 = time.perf_counter()
    tuned = batch.tune([filters.FilterSet.parse(text) for _, text, _ in mix])
    tuned["walk_s"] = time.perf_counter() - t0
    progress(f"slot engine walk: {tuned}")
    batched = run_mix(mix, seconds, tuned["slots"], tuned["chains"], tuned["half"],
                      batch.LEAST_SERVED)
    progress("batched mix done")
    plain = run_mix(mix, seconds, 1, tuned["chains"], tuned["half"], batch.ROUND_ROBIN)
    full = run_mix(mix, seconds, 1, 1 << 18, 256, batch.ROUND_ROBIN, ring_cap=1 << 21)
    progress("one search per launch done")
    return {
        "workload": "64 searches (search_mix seed 4: " + ", ".join(f"{v} {k}" for k, v in
                                                                   sorted(kinds.items())) +
                    "; wanted hits log-uniform 1..1000), fixed wall-time window",
        "unit_shape": tuned,
        "batched": batched,
        "plain_per_search_shape": plain,
        "plain_full_gpu": full,
        "speedup_vs_plain": batched["keys_per_s"] / plain["keys_per_s"],
        "speedup_vs_plain_full_gpu": batched["keys_per_s"] / full["keys_per_s"],
    }


def _family(text):
    return text.split(":")[0] if ":" in text else "prefix"


def served_leg(seconds, rate=200.0, fifo_seconds=60.0, unit=(64, 16384, 256)):
    """Poisson arrivals at `rate`/s for `seconds`, each request a filter of
    the batched mix wanting 1, 10 or 100 hits, through the continuous-batching
    loop.  The comparator is like for like: the requests the loop completed,
    all available at once, served one at a time on the whole GPU (first come,
    first served) for at most `fifo_seconds`; both are reported as requests
    completed per second and as candidates per second per filter family."""
    from onion_gen_amd import batch, filters, serve, workload
    arrivals = workload.poisson_arrivals(rate, seconds, 5)
    mix = workload.search_mix(len(arrivals), 6, hits="choice")
    reqs = [(t, seed, text, want) for t, (seed, text, want) in zip(arrivals, mix)]
    keep = ("wall_s", "keys_per_s", "completed", "running", "queued", "rejected",
            "max_resident", "p50_time_to_n", "p99_time_to_n", "p50_time_to_first",
            "p99_time_to_first")
    st, res = serve.run(reqs, duration=seconds, slots=unit[0], chains=unit[1], half=unit[2],
                        max_queue=1 << 16, policy=batch.LEAST_SERVED)
    progress(f"served, continuous: {st['completed']} of {len(reqs)} completed")
    cont = {k: st[k] for k in keep}
    done = [i for i, r in enumerate(res) if r["status"] == "done"]
    cont["completed_per_s"] = len(done) / st["wall_s"]

    def by_family(idx, results, wall_of):
        fam = {}
        for i in idx:
            f = fam.setdefault(_family(reqs[i][2]), {"requests": 0, "examined": 0, "done": 0})
            f["requests"] += 1
            f["examined"] += results[i]["examined"]
            f["done"] += results[i]["status"] == "done"
        for name, f in fam.items():
            f["keys_per_s"] = f["examined"] / wall_of(name) if wall_of(name) else None
        return fam

    # continuous: every family runs the whole window side by side
    cont["families"] = by_family(range(len(reqs)), res, lambda name: st["wall_s"])
    # the same completed requests, one at a time on the whole GPU
    same = [(0.0, reqs[i][1], reqs[i][2], reqs[i][3]) for i in done]
    with batch.Batch(slots=1, chains=1 << 18, half=256) as b:
        one = max(b.footprint(filters.FilterSet.parse(t)) for t in {r[2] for r in same})
    st2, res2 = serve.run(same, duration=fifo_seconds, slots=1, chains=1 << 18, half=256,
                          max_queue=1 << 16, device_budget=one, policy=batch.ROUND_ROBIN)
    progress(f"served, one at a time: {st2['completed']} of {len(same)} completed")
    fifo = {k: st2[k] for k in keep}
    fifo["completed_per_s"] = st2["completed"] / st2["wall_s"]
    # one at a time: a family's rate is over the time its requests held the GPU
    held = {}
    for j, r in enumerate(res2):
        if r["admitted_s"] >= 0:
            end = r["done_s"] if r["done_s"] >= 0 else st2["wall_s"]
            name = _family(same[j][2])
            held[name] = held.get(name, 0.0) + end - r["admitted_s"]
    fam2 = {}
    for j, r in enumerate(res2):
        f = fam2.setdefault(_family(same[j][2]), {"requests": 0, "examined": 0, "done": 0})
        f["requests"] += 1
        f["examined"] += r["examined"]
        f["done"] += r["status"] == "done"
    for name, f in fam2.items():
        f["keys_per_s"] = f["examined"] / held[name] if held.get(name) else None
    fifo["families"] = fam2
    return {
        "workload": f"Poisson {rate:.0f} req/s for {seconds:.0f} s ({len(reqs)} requests), "
                    "filters of the batched mix, n in {1, 10, 100}",
        "unit_shape": {"slots": unit[0], "chains": unit[1], "half": unit[2]},
        "continuous": cont,
        "fifo_same_completed": fifo,
        "fifo_workload": f"the {len(same)} requests the continuous loop completed, all at t = 0, "
                         f"one at a time on 2^18 chains, at most {fifo_seconds:.0f} s",
        "completed_per_s_vs_fifo": (cont["completed_per_s"] / fifo["completed_per_s"]
                                    if fifo["completed_per_s"] else None),
        "keys_per_s_vs_fifo": (cont["keys_per_s"] / fifo["keys_per_s"]
                               if fifo["keys_per_s"] else None),
    }


def dictionary_leg(threads, half, rounds):
    """The dictionary workload: 10,000 distinct prefixes of 5-8 symbols
    (ChaCha20 seeded 2) behind a 24-bit prefix bitmap and the exact table, 2^34
    candidates, then every survivor through the batched hit sink (secret at the
    offset, public key re-derived on the device, hostname).  Survivors stay in
    one ctypes buffer from the engine to the sink."""
    import ctypes
    import numpy as np
    from onion_gen_amd import filters, gpu, key, workload
    from onion_gen_amd._lib import OgError, err_buf, lib
    from onion_gen_amd.gpu import _Survivor
    from onion_gen_amd.key import _KeyRecord
    seed = workload.chacha_stream(0, 32)
    fs = filters.FilterSet.parse(workload.dictionary(10000, 2), index_bits=24)
    eng = gpu.Engine.with_filters(key.start_point(seed), fs, threads=threads, half=half,
                                  rounds=rounds)
    batch = eng.batch()
    steps = -(-(1 << 34) // batch)
    room = eng.room()
    buf = (_Survivor * (steps * room))()
    base = ctypes.addressof(buf)
    n = 0
    e = err_buf()
    cnt = ctypes.c_uint32()
    t0 = time.perf_counter()
    for _ in range(steps):
        if lib().og_hip_engine_advance(eng._h, ctypes.c_void_p(base + n * ctypes.sizeof(_Survivor)),
                                       room, ctypes.byref(cnt), e, len(e)) != 0:
            raise OgError(e.value.decode())
        n += cnt.value
    search_s = time.perf_counter() - t0
    eng.close()
    recs = (_KeyRecord * max(n, 1))()
    t1 = time.perf_counter()
    if lib().og_hip_finish_hits(0, seed, buf, n, recs, e, len(e)) != 0:
        raise OgError(e.value.decode())
    sink_s = time.perf_counter() - t1
    raw = np.frombuffer(recs, dtype=np.uint8).reshape(-1, ctypes.sizeof(_KeyRecord))[:n]
    status = raw[:, _KeyRecord.status.offset:_KeyRecord.status.offset + 4].copy().view(np.int32)
    return {
        "workload": "10,000 prefixes of 5-8 symbols (ChaCha20 seeded 2), 24-bit bitmap + exact "
                    f"table, {steps * batch} candidates ({steps} launches of {batch})",
        "keys_per_s": steps * batch / search_s,
        "survivors": n,
        "sink_s": sink_s,
        "sink_hits_per_s": n / sink_s if sink_s > 0 else None,
        "verified": int((status == 0).sum()),
        "clamping_dropped": int((status == 1).sum()),
        "mismatched": int((status == 2).sum()),
    }


FORMS = [
    ("suffix:2xaqd", "a 5-symbol suffix: every candidate needs sign and checksum"),
    ("contains:q2x7a", "a 5-symbol substring"),
    ("regex:^ab[2-7]cd", "an anchored expression, settled by the key"),
    ("regex:(cat|dog)food", "an unanchored expression, two obligatory sequences"),
    ("regex:^zz.*qd$", "an expression reaching the checksum"),
    ("dict:contains8", "1,000 eight-symbol substrings (Python random, seed 5): one automaton"),
    ("dict:contains5", "1,000 five-symbol substrings (Python random, seed 5): one automaton, "
                       "most candidates hashed (a match can still end in the checksum)"),
    ("dict:contains8:10000", "10,000 eight-symbol substrings (Python random, seed 5): two "
                             "automata"),
]


def form_lines(text):
    """The filter lines of a forms-leg entry."""
    if not text.startswith("dict:"):
        return [text]
    import random
    from onion_gen_amd import workload
    rng = random.Random(5)
    spec = text[len("dict:contains"):].split(":")
    length, count = int(spec[0]), int(spec[1]) if len(spec) > 1 else 1000
    return ["contains:" + "".join(rng.choice(workload.ALPHABET) for _ in range(length))
            for _ in range(count)]


def forms_leg(threads, half, rounds, launches):
    """The other filter forms on the headline shape: `launches` timed launches
    after one warm-up, each form in its own engine; every survivor is a
    verified match of the form (the automata decide expressions exactly)."""
    import ctypes
    from onion_gen_amd import filters, gpu, key, workload
    from onion_gen_amd._lib import OgError, err_buf, lib
    from onion_gen_amd.gpu import _Survivor
    seed = workload.chacha_stream(0, 32)
    start = key.start_point(seed)
    out = {}
    e = err_buf()
    cnt = ctypes.c_uint32()
    for text, what in FORMS:
        fs = filters.FilterSet.parse(form_lines(text))
        eng = gpu.Engine.with_filters(start, fs, threads=threads, half=half, rounds=rounds)
        batch, room = eng.batch(), eng.room()
        buf = (_Survivor * room)()  # survivors stay in one ctypes buffer

        def advance():
            if lib().og_hip_engine_advance(eng._h, buf, room, ctypes.byref(cnt), e, len(e)) != 0:
                raise OgError(e.value.decode())
            return cnt.value
        advance()
        hits, kernel_ms = 0, 0.0
        t0 = time.perf_counter()
        for _ in range(launches):
            hits += advance()
            kernel_ms += eng.last_ms()
        wall = time.perf_counter() - t0
        eng.close()
        # wall includes copying the survivors back; the kernel rate does not
        out[text] = {"what": what, "keys_per_s": launches * batch / wall,
                     "kernel_keys_per_s": launches * batch / (kernel_ms * 1e-3),
                     "kernel_avg_ms": kernel_ms / launches, "hits": hits,
                     "candidates": launches * batch}
    return out


def host_cpus():
    """(threads to use, description): the CPUs this process may run on, at most
    the CPU quota of its control group (a container's share of a larger host),
    with the processor model and the host's count beside them."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = min(usable, quota) if quota else usable
    return threads, {"cpu_model": model, "host_cpus": os.cpu_count(), "affinity": usable,
                     "cgroup_quota": quota}


def cpu_baseline(prefix, seconds):
    """The reference's CPU engines restated in C (the oracle), built for this
    host (-O3 -march=native) where it is timed, on every CPU this process may
    use, one block per thread: PairEngine at its default batch of 8192
    (src/pairs.rs, the reference's default and fastest CPU path, with its limb
    probe) and BatchEngine at width 256 (src/batch.rs), each the median of
    three runs of `seconds` / 3 (BatchEngine: half that)."""
    import ctypes
    import statistics
    import subprocess
    import tempfile
    from tests.oracle_lib import load
    from onion_gen_amd import workload
    out_dir = tempfile.mkdtemp(prefix="og_oracle_native_")
    lib_path = os.path.join(out_dir, "liboracle_native.so")
    flags = "-O3 -march=native"
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                        f"NATIVE_OUT={lib_path}"], check=True, capture_output=True)
        L = load(lib_path)
    except (subprocess.CalledProcessError, OSError):
        L, flags = load(), "-O2 -march=x86-64-v2 (the native build failed)"
    cores, cpu = host_cpus()
    err = ctypes.create_string_buffer(256)
    fs = L.ofilter_parse(prefix.encode(), err, 256)
    seed = workload.chacha_stream(0, 32)
    ex = ctypes.c_uint64()
    pairs, batch, n_pairs, n_batch = [], [], 0, 0
    for _ in range(3):
        pairs.append(L.obench_pairs(seed, 8192, cores, seconds / 3, fs, ctypes.byref(ex)))
        n_pairs += ex.value
    for _ in range(3):
        batch.append(L.obench_batch(seed, 256, cores, seconds / 6, fs, ctypes.byref(ex)))
        n_batch += ex.value
    L.ofilter_free(fs)
    return {"value": statistics.median(pairs), "unit": "keys/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle PairEngine batch 8192 (src/pairs.rs restated in C, limb probe), "
                      f"built {flags} on this host, one block per thread on {cores} threads, "
                      f"prefix '{prefix}', median of 3 runs of {seconds / 3:.1f} s "
                      f"({n_pairs} candidates in all)",
            "runs": pairs, "build": flags, **cpu,
            "batch_engine": {"value": statistics.median(batch), "runs": batch,
                             "sample": f"oracle BatchEngine width 256 (src/batch.rs), median of 3 "
                                       f"runs of {seconds / 6:.1f} s ({n_batch} candidates)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="0: cover 2^34 candidates")
    ap.add_argument("--warmup", type=int, default=2)
    # The shape the autotune walk (og_hip_engine_tune) picks for this workload on
    # MI355X: 2^20 chains, 1024 pairs per window, one window per launch (86 GB
    # of scratch planes: two 40-byte elements per pair).
    ap.add_argument("--threads", type=int, default=1 << 20)
    ap.add_argument("--half", type=int, default=1024)
    ap.add_argument("--rounds", type=int, default=1)
    ap.add_argument("--no-dict", action="store_true",
                    help="skip the dictionary leg (10,000 prefixes over 2^34 candidates, every "
                         "hit through the hit sink)")
    ap.add_argument("--tuned", action="store_true",
                    help="run the autotune walk first and search with the shape it picks")
    ap.add_argument("--cpu-seconds", type=float, default=18.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--filter", default="",
                    help="search this filter instead of the 6-symbol prefix (for profiling the "
                         "other forms; the bench line then names it; dict:contains<L>[:<count>] "
                         "names a substring dictionary of the forms leg)")
    ap.add_argument("--mix-seconds", type=float, default=5.0,
                    help="wall-time window of each batched-mix run (0 skips the leg)")
    ap.add_argument("--serve-seconds", type=float, default=120.0,
                    help="arrival window of the continuous served run (0 skips the leg)")
    ap.add_argument("--fifo-seconds", type=float, default=60.0,
                    help="limit of the one-at-a-time replay of the completed requests")
    ap.add_argument("--forms-launches", type=int, default=4,
                    help="timed launches per filter form in the forms leg (0 skips it)")
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
    text = args.filter or prefix
    start = key.start_point(seed)
    fs = filters.FilterSet.parse(form_lines(text))
    tuned = None
    if args.tuned:
        tuned = gpu.Engine.tune(start, fs)
        args.threads, args.half, args.rounds = tuned["threads"], tuned["half"], tuned["rounds"]
    p_mul, p_clock = measured_mad_peak()
    eng = gpu.Engine.with_filters(start, fs, threads=args.threads, half=args.half,
                                  rounds=args.rounds, hit_cap=0 if not args.filter else 1 << 22)
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
    progress(f"headline: {steps} launches, {rate:.4g} keys/s")

    # every survivor through the hit sink: its secret at the offset, its
    # public key re-derived on the device from the scalar alone and compared
    # with the survivor, then the whole address matched on the host
    # (a filter of the forms, run for profiling, may return millions: there
    # an evenly spaced sample of 65,536 of them, which the line states)
    sample = survivors
    if args.filter and len(survivors) > 65536:
        sample = survivors[::-(-len(survivors) // 65536)]
    recs = key.finish_hits(seed, sample)
    verified = sum(1 for r in recs if r.status == key.HIT_OK and
                   fs.match_address(key.address_bytes(r.public_key)) >= 0)
    clamping = sum(1 for r in recs if r.status == key.HIT_CLAMPING)

    F = field_muls_per_candidate(args.half, key_form=eng.kernels() == 1,
                                 with_x=text.startswith("suffix:"))
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
        "config": {"workload": (f"one 6-symbol prefix '{prefix}', {examined} sequential "
                                f"candidates (2^34 at the default step count)") if not args.filter
                   else f"one filter '{text}', {examined} sequential candidates",
                   "threads": args.threads, "half": args.half, "rounds": args.rounds,
                   "candidates_per_step": batch, "tuning_walk": tuned},
        "hits": len(survivors),
        "hits_rederived": verified,  # through the hit sink
        "hits_checked": len(sample),
        "hits_clamping_dropped": clamping,
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": p_mul / 1e12,
            "unit": "T mad32/s",
            "frac": achieved / p_mul,
            "peak_source": (f"measured: v_mad_u64_u32 issued back to back by og_hip_issue_bench "
                            f"at a shader clock of {p_clock / 1e9:.3f} GHz"),
            "peak_nominal": VALU_PEAK_LANE_OPS / 1e12,
            "frac_nominal": achieved / VALU_PEAK_LANE_OPS,
            "traffic": None if args.filter else measured_traffic(args.threads, args.half,
                                                                  args.rounds),
            "traffic_source": "profiles/search_prefix6_traffic.md (rocprofv3 FETCH_SIZE x2.0 "
                              "+ WRITE_SIZE x1.0, copy-calibrated)",
            "kernel_avg_ms": kernel_ms / steps,
            "field_muls_per_candidate": F,
        },
    }
    eng.close()
    if not args.no_dict and not args.filter:
        out["dictionary"] = dictionary_leg(args.threads, args.half, args.rounds)
        progress("dictionary leg done")
    if args.forms_launches > 0 and not args.filter:
        out["forms"] = forms_leg(args.threads, args.half, args.rounds, args.forms_launches)
        progress("forms leg done")
    if args.mix_seconds > 0:
        out["batched"] = batched_leg(args.mix_seconds)
    if args.serve_seconds > 0:
        unit = (64, 16384, 256)
        if "batched" in out:
            u = out["batched"]["unit_shape"]
            unit = (u["slots"], u["chains"], u["half"])
        out["served"] = served_leg(args.serve_seconds, fifo_seconds=args.fifo_seconds, unit=unit)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prefix, args.cpu_seconds)
        progress("cpu baseline done")
    if verified + clamping != len(sample):
        out["error"] = (f"{len(sample) - verified - clamping} of {len(sample)} survivors checked "
                        "did not re-derive to a matching key")
    print(json.dumps(out), flush=True)
    if "error" in out:
        sys.exit(1)


if __name__ == "__main__":
    main()
