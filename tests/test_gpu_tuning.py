"""The autotune walk (og_hip_engine_tune) and the tuned constructor: the walk
records its fastest shape under the device's key, an engine created with
threads = 0 runs that shape, and its survivors are real hits."""
import os

import pytest

from onion_gen_amd import filters, gpu, key, tuning, workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_walk_records_and_the_tuned_engine_uses_it(tmp_path, monkeypatch):
    monkeypatch.setenv("OG_TUNING_PATH", str(tmp_path / "tuning.yaml"))
    say("case: tuning walk over 20 shapes, 4-symbol prefix")
    seed = workload.chacha_stream(0, 32)
    start = key.start_point(seed)
    fs = filters.FilterSet.parse("2xeq")
    best = gpu.Engine.tune(start, fs)
    say(f"best: {best}")
    assert best["threads"] in (1 << 20, 1 << 19, 1 << 18, 1 << 17, 1 << 16)
    assert best["half"] in (1024, 512, 256, 128) and best["rounds"] >= 1
    assert best["rate"] > 1e9
    dev = gpu.look()[0]
    cached = tuning.get(f"device/{dev.name.replace('/', '_')}/hip/0/1.0.2")
    assert cached and cached["second"] == best["threads"] and cached["first"] == best["half"]
    with gpu.Engine.with_filters(start, fs) as eng:
        assert eng.shape() == (best["threads"], best["half"], best["rounds"])
        hits = eng.advance()
        assert hits
        for s in hits[:32]:
            pk, addr = key.finish_hit(seed, s.offset)
            assert addr.startswith("2xeq") and pk[:31] == s.packed[:31]
    # the reference's constructor with threads = 0 takes the same cached shape
    words, bits = filters.FilterSet.parse(workload.dictionary(1000, seed=3)).index_words()
    with gpu.Engine.new(start, 0, words, bits) as eng:
        assert eng.shape() == (best["threads"], best["half"], best["rounds"])


def test_tuned_engine_fits_the_memory_budget(tmp_path, monkeypatch):
    """threads = 0 under a 32 GiB budget: the walk builds no shape above it,
    and the engine it makes holds at most that much device memory, as the
    reference's walk is bounded by memory (src/gpu/mod.rs:314-342)."""
    monkeypatch.setenv("OG_TUNING_PATH", str(tmp_path / "tuning.yaml"))
    budget = 32 << 30
    seed = workload.chacha_stream(0, 32)
    start = key.start_point(seed)
    fs = filters.FilterSet.parse("2xeq")
    say("case: tuning walk under a 32 GiB budget, 4-symbol prefix")
    best = gpu.Engine.tune(start, fs, max_bytes=budget)
    say(f"best: {best}")
    with gpu.Engine.with_filters(start, fs, max_bytes=budget) as eng:
        say(f"engine: {eng.shape()} holds {eng.bytes()} bytes")
        assert eng.shape() == (best["threads"], best["half"], best["rounds"])
        assert 0 < eng.bytes() <= budget
        assert eng.advance()
    # a remembered shape above a smaller budget is measured again
    with gpu.Engine.with_filters(start, fs, max_bytes=8 << 30) as eng:
        say(f"engine under 8 GiB: {eng.shape()} holds {eng.bytes()} bytes")
        assert 0 < eng.bytes() <= 8 << 30


def test_batch_walk_records_and_the_zero_shape_uses_it(tmp_path, monkeypatch):
    """og_batch_tune times the slot engine's unit shapes with the 64-search
    mix's filters under a 48 GiB budget and the default 25 ms step bound,
    records the pick under the device's hip-batch key, and og_batch_create
    with slots, chains and half all 0 runs that shape (src/tuning.rs:27-50 for
    the cache, src/gpu/mod.rs:346-428 for the walk)."""
    from onion_gen_amd import batch
    monkeypatch.setenv("OG_TUNING_PATH", str(tmp_path / "tuning.yaml"))
    mix = workload.search_mix(64, 4)
    sets = [filters.FilterSet.parse(text) for _, text, _ in mix]
    say("case: slot engine tuning walk, the 64-search mix, 48 GiB")
    best = batch.tune(sets, max_bytes=48 << 30)
    say(f"best: {best}")
    assert best["slots"] in (32, 64, 128) and best["chains"] in (4096, 8192, 16384, 32768)
    assert best["half"] in (256, 512) and best["rate"] > 1e9
    # only shapes whose steps average at most 25 ms take part (every slot
    # busy, so a step examines the whole unit)
    assert best["step_ms"] <= 25.0
    dev = gpu.look()[0]
    cached = tuning.get(tuning.device_key(dev.name, "hip-batch", 0))
    assert cached and cached["slots"] == best["slots"] and cached["second"] == best["chains"]
    assert cached["first"] == best["half"]
    with batch.Batch(slots=0, chains=0, half=0) as b:
        assert b.per_step() == best["chains"] * (2 * best["half"] + 1)
        seed, text, _ = mix[0]
        sid = b.admit(key.start_point(seed), filters.FilterSet.parse(text), 0)
        b.set_want(sid, 10)  # alone it widens to every slot: its hits would fill the ring
        b.step()
        assert b.search_state(sid)["examined"] >= b.per_step()
