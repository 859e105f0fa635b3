"""The batched hit sink against independent derivations: every survivor of a
search becomes a verified key whose secret produces its public key, clamping
and mismatch are reported, and the records write tor's key files."""
import hashlib
import os

import pytest

from onion_gen_amd import filters, gpu, key, workload
from onion_gen_amd.gpu import Survivor
from tests import ed25519_ref

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_survivors_become_verified_keys(tmp_path):
    say("case: 2-symbol prefix, 2^20 candidates, finish every survivor")
    seed = workload.chacha_stream(0, 32)
    eng = gpu.Engine.with_filters(key.start_point(seed), filters.FilterSet.parse("2x"),
                                  threads=4096, half=128)
    survivors = []
    while len(survivors) < 1000:
        survivors += eng.advance()
    eng.close()
    recs = key.finish_hits(seed, survivors)
    assert len(recs) == len(survivors)
    for s, r in zip(survivors, recs):
        assert r.status == key.HIT_OK and r.offset == s.offset
        # the host's own scalar multiplication and key derivation agree
        assert r.expanded_secret == key.expanded_secret_at_offset(seed, s.offset)
        assert r.public_key == key.public_from_scalar(r.expanded_secret[:32])
        assert r.public_key[:31] == s.packed[:31]
        assert r.hostname == key.address(r.public_key) + ".onion"
        assert r.hostname.startswith("2x")
    # a sample by textbook big-integer arithmetic
    for r in recs[:20]:
        assert ed25519_ref.encode(ed25519_ref.mul(int.from_bytes(r.expanded_secret[:32],
                                                                  "little"))) == r.public_key
    created, path = key.write_key(tmp_path, recs[0].public_key, recs[0].expanded_secret)
    assert created and path.endswith(recs[0].hostname)


def test_clamping_and_mismatch_are_reported():
    say("case: hit sink status codes")
    seed = hashlib.sha256(b"sink").digest()
    good = key.finish_hit(seed, 8 * 77)
    y = good[0][:31] + bytes([good[0][31] & 0x7F])
    bad_y = bytes([y[0] ^ 1]) + y[1:]
    recs = key.finish_hits(seed, [Survivor(8 * 77, y), Survivor(8 * 77 + 1, y),
                                  Survivor(8 * 77, bad_y)])
    assert [r.status for r in recs] == [key.HIT_OK, key.HIT_CLAMPING, key.HIT_MISMATCH]
    assert recs[0].public_key == good[0]


def test_random_scalars_match_the_host_derivation():
    """The device comb over 4096 secrets at random offsets, every window value
    exercised, against the host's double-and-add."""
    say("case: hit sink on 4096 random offsets")
    import random
    rng = random.Random(7)
    seed = hashlib.sha256(b"comb").digest()
    offs = [8 * rng.getrandbits(40) for _ in range(4096)]
    surv = []
    for o in offs:
        pk = key.finish_hit(seed, o)[0]
        surv.append(Survivor(o, pk[:31] + bytes([pk[31] & 0x7F])))
    recs = key.finish_hits(seed, surv)
    assert all(r.status == key.HIT_OK for r in recs)
    assert all(r.public_key == key.finish_hit(seed, o)[0] for r, o in zip(recs, offs))
