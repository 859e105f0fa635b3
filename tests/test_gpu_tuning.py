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
