"""The lease adapter: a proto::Lease searched block by block on the device.
Every block's finds are exactly the oracle's hits over the candidates it
examined from block_seed(root, block), and the master-side derivation of each
find (secret at (block, offset)) gives an address the filter matches."""
import hashlib

import pytest

from onion_gen_amd import filters, gpu, key, lease
from tests.oracle_lib import OracleFilters, search

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def check(lz, lines, found, per_block):
    root = bytes.fromhex(lz["seed"])
    by_block = {}
    for f in found:
        assert f["lease"] == lz["lease"]
        by_block.setdefault(f["block"], set()).add(f["offset"] // 8)
        secret = key.expanded_secret_at_offset(key.block_seed(root, f["block"]), f["offset"])
        pk = key.public_from_scalar(secret[:32])
        assert OracleFilters(lines).match_address(key.address_bytes(pk)) >= 0
    for b in range(lz["first_block"], lz["first_block"] + lz["blocks"]):
        seed = key.block_seed(root, b)
        want = {k for k, _, _ in search(key.start_point(seed), 0, per_block, 256,
                                        OracleFilters(lines))}
        assert by_block.get(b, set()) == want and want, b


def test_lease_blocks_match_the_oracle():
    say("case: lease of 3 blocks from block 400, 4 launches each, prefix 2x")
    lz = {"lease": 7, "seed": hashlib.sha256(b"root").hexdigest(), "first_block": 400,
          "blocks": 3, "expires_in": 60}
    fs = filters.FilterSet.parse("2x")
    with gpu.Engine.with_filters(key.start_point(bytes(32)), fs, threads=1024, half=64) as eng:
        res = lease.run(lz, fs, eng, launches=4)
        per_block = 4 * eng.batch()
    assert res["heartbeat"] == {"lease": 7, "blocks_done": 3, "candidates": 3 * per_block}
    assert res["blocks_per_second"] > 0
    check(lz, ["2x"], res["found"], per_block)


def test_bitmap_only_engine_is_confirmed_by_the_set():
    say("case: lease on the reference constructor (bitmap only), dictionary of 3")
    lines = ["2x", "q7", "abc"]
    fs = filters.FilterSet.parse(lines)
    words, bits = fs.index_words()
    lz = {"lease": 8, "seed": hashlib.sha256(b"root2").hexdigest(), "first_block": 0,
          "blocks": 2, "expires_in": 60}
    with gpu.Engine.new(key.start_point(bytes(32)), 1024, words, bits) as eng:
        res = lease.run(lz, fs, eng, launches=2)
        per_block = 2 * eng.batch()
    check(lz, lines, res["found"], per_block)
