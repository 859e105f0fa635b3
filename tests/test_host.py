"""CPU tests of the library's host side: exports, key derivation, filter sets.

The product's host code (onion_gen_amd/csrc/og_host.cpp, og_filters.cpp) is a
second restatement of src/key.rs and src/filter.rs; the oracle is the checker.
"""
import hashlib
import os
import random
import re

import pytest

from onion_gen_amd import _lib, filters, gpu, key, workload
from tests.oracle_lib import OracleFilters, address as o_address, public_key as o_public_key, \
    start_point as o_start_point

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "onion_gen_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(og_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_no_device_is_an_error_not_a_fallback():
    if gpu.device_count() > 0:
        pytest.skip("a device is visible")
    fs = filters.FilterSet.parse("abc")
    with pytest.raises(_lib.OgError, match="no HIP device"):
        gpu.Engine.with_filters(key.start_point(bytes(32)), fs)


def test_key_derivation_matches_oracle():
    rng = random.Random(31)
    for i in range(30):
        seed = bytes(rng.getrandbits(8) for _ in range(32))
        assert key.start_point(seed) == o_start_point(seed), i
        pk = key.public_from_scalar(key.secret_scalar(seed))
        assert pk == o_public_key(seed), i
        assert key.address(pk) == o_address(pk), i
        assert key.secret_scalar(seed) == key.expanded_secret(seed)[:32]
    root = bytes(range(32))
    assert key.block_seed(root, 7) == hashlib.sha512(
        b"onion-gen block v1" + root + (7).to_bytes(8, "little")).digest()[:32]


def test_offset_secret_rederives_the_chain_key():
    seed = bytes(range(32))
    for k in (0, 1, 7, 2047, 123456):
        sk = key.expanded_secret_at_offset(seed, 8 * k)
        assert sk is not None
        s0 = int.from_bytes(key.secret_scalar(seed), "little")
        assert int.from_bytes(sk[:32], "little") == s0 + 8 * k
        pk, addr = key.finish_hit(seed, 8 * k)
        assert len(addr) == 56 and addr.endswith("d")


def host_field_op(op, a, b):
    import ctypes
    n = len(a)
    out = ctypes.create_string_buffer(32 * n)
    rc = _lib.lib().og_field_op_host(op, b"".join(x.to_bytes(32, "little") for x in a),
                                     b"".join(x.to_bytes(32, "little") for x in b), out, n)
    assert rc == 0
    return [int.from_bytes(out.raw[32 * i:32 * i + 32], "little") for i in range(n)]


def test_field_arithmetic_both_forms_on_the_host():
    """The device field code (8 x 32 and 10 x 25.5) compiled for the host,
    against Python integers, including the loosest operands the search feeds
    its multiplications (op 16)."""
    P = 2 ** 255 - 19
    rng = random.Random(17)
    edges = [0, 1, 2, 19, P - 1, P, P + 1, 2 * P, 2 ** 255 - 1, 2 ** 255, 2 ** 256 - 1,
             2 ** 256 - 38, 2 ** 256 - 39, 38, 2 ** 26 - 1, 2 ** 26, 2 ** 51]
    a = edges + [rng.getrandbits(256) for _ in range(2000)]
    b = list(reversed(edges)) + [rng.getrandbits(256) for _ in range(2000)]
    for base in (0, 10):
        assert host_field_op(base + 0, a, b) == [x * y % P for x, y in zip(a, b)]
        assert host_field_op(base + 1, a, b) == [x * x % P for x in a]
        assert host_field_op(base + 2, a, b) == [(x + y) % P for x, y in zip(a, b)]
        assert host_field_op(base + 3, a, b) == [(x - y) % P for x, y in zip(a, b)]
        assert host_field_op(base + 5, a, b) == [x % P for x in a]
        nz = [x for x in a if x % P][:150]
        assert host_field_op(base + 4, nz, nz) == [pow(x, P - 2, P) for x in nz]
    assert host_field_op(16, a, b) == [(x - y) * (2 * x - y) % P for x, y in zip(a, b)]
    assert host_field_op(25, a, b) == [(x - y) * y % P for x, y in zip(a, b)]
    # the divstep inversion, against Fermat, over the edges and random values
    # (0 maps to 0), and on values in [p, 2^255) as a product leaves them
    inv_in = a[:167] + [P, P + 1, P + 18, 2 ** 255 - 1, 2 ** 255 + P, 2 ** 256 - 1]
    assert host_field_op(26, inv_in, inv_in) == [pow(x % P, P - 2, P) for x in inv_in]
    # a product's output canonicalised without a carry pass first: values in
    # [p, 2^255) kept as they stand (op 17), and products (op 18)
    above = [P, P + 1, P + 18, 2 ** 255 - 1, 2 ** 255 + P, 2 ** 256 - 1]
    a17 = edges + above + [rng.getrandbits(256) for _ in range(500)]
    assert host_field_op(17, a17, a17) == [(x % 2 ** 255) % P for x in a17]
    assert host_field_op(18, a, b) == [x * y % P for x, y in zip(a, b)]
    assert host_field_op(19, a, b) == [(x % 2 ** 255) * y % P & 1 for x, y in zip(a, b)]
    # products that land on [p, 2^255) before reduction: x times one
    assert host_field_op(19, above, [1] * len(above)) == [(x % 2 ** 255) % P & 1 for x in above]
    # products by a constant (the operand of element i & ~63): one product and a
    # double product with one carry pass, as the search loop takes them
    lead_a = [a[i & ~63] for i in range(len(a))]
    lead_b = [b[i & ~63] for i in range(len(b))]
    assert host_field_op(21, a, b) == [x * c % P for x, c in zip(a, lead_b)]
    assert host_field_op(22, a, b) == [(x * c + y * e) % P
                                       for x, y, c, e in zip(a, b, lead_b, lead_a)]
    # the key forms' probe word (low 32 bits) and its exactness flag (bit 32):
    # where exact it is the canonical low word of (2x - y)(x - y), else that
    # minus it is 0, 19 or 38; rarely inexact
    got = host_field_op(23, a, b)
    diffs = [((2 * x - y) * (x - y) % P - w) % 2 ** 32 for x, y, w in zip(a, b, got)]
    exact = [w >> 32 == 1 for w in got]
    assert all(d == 0 for d, e in zip(diffs, exact) if e)
    assert all(d in (0, 19, 38) for d in diffs)
    assert exact[len(edges):].count(False) <= len(a) // 200  # random operands: about 1 in 1,024
    # the three-column word (op 27): bits 0..25 of the same, on the same terms,
    # and op 23's word less column 1
    got26 = host_field_op(27, a, b)
    d26 = [((2 * x - y) * (x - y) % P - w) % 2 ** 26 for x, y, w in zip(a, b, got26)]
    assert [w >> 32 == 1 for w in got26] == exact
    assert all(d == 0 for d, e in zip(d26, exact) if e)
    assert all(d in (0, 19, 38) for d in d26)
    assert [w & (2 ** 26 - 1) for w in got26] == [w % 2 ** 26 for w in got]
    assert all(w % 2 ** 32 < 2 ** 26 for w in got26)
    # the sign of x as the search takes it: from three columns where they settle it
    assert host_field_op(24, a, b) == [(x % 2 ** 255) * y % P & 1 for x, y in zip(a, b)]
    assert host_field_op(24, above, [1] * len(above)) == [(x % 2 ** 255) % P & 1 for x in above]
    # the largest limbs a tight value can carry: all-ones limbs
    big = [sum(((1 << (26 if i % 2 == 0 else 25)) - 1) << [0, 26, 51, 77, 102, 128, 153, 179, 204, 230][i]
               for i in range(10))]
    assert host_field_op(16, big * 3, [0, 1, big[0]]) == \
        [(big[0] - y) * (2 * big[0] - y) % P for y in [0, 1, big[0]]]


def test_hit_room_bounds():
    # src/gpu/mod.rs:521-553
    for bits in range(8, 31):
        for threads in (1 << 12, 1 << 16, 1 << 17):
            batch = threads * 64
            assert gpu.hit_room(batch, bits) >= (batch >> bits) * 4
    assert gpu.hit_room(2 ** 64 - 1, 0) * 40 <= 16 * 1024 * 1024
    assert gpu.hit_room(1 << 22, 30) >= 1024


def test_filter_sets_match_oracle_structure():
    cases = [
        ["abc"],
        ["abc", "abcd", "prefix:xyz", "suffix:ad", "contains:qq[a-c]"],
        workload.dictionary(1000, seed=3),
        workload.dictionary(10000, seed=2),
    ]
    for lines in cases:
        mine = filters.FilterSet.parse(lines)
        theirs = OracleFilters(lines)
        assert len(mine) == len(theirs)
        a, b = mine.index_words(), theirs.index()
        assert (a is None) == (b is None)
        if a:
            assert a[1] == b[1] and a[0] == b[0]
    fixed = filters.FilterSet.parse(workload.dictionary(10000, seed=2), index_bits=24)
    assert fixed.index_words()[1] == 24
    assert fixed.index_words()[0] == OracleFilters(workload.dictionary(10000, seed=2)).index(24)[0]


def test_filter_matching_matches_oracle():
    words = workload.dictionary(3000, seed=9)
    general = ["suffix:[a-z]qd", "contains:a2?c", "prefix:ab?d", "suffix:aid"]
    mine_p, theirs_p = filters.FilterSet.parse(words), OracleFilters(words)
    mine_g, theirs_g = filters.FilterSet.parse(general), OracleFilters(general)
    rng = random.Random(9)
    for i in range(3000):
        pk = bytearray(rng.getrandbits(8) for _ in range(32))
        if i % 2:
            w = words[rng.randrange(len(words))]
            v = int("".join(format(workload.ALPHABET.index(c), "05b") for c in w), 2)
            nb = 5 * len(w)
            h = int.from_bytes(pk[:8], "big")
            pk[:8] = ((h & ((1 << (64 - nb)) - 1)) | (v << (64 - nb))).to_bytes(8, "big")
        pk = bytes(pk)
        ab = key.address_bytes(pk)
        assert ab == pk + hashlib.sha3_256(b".onion checksum" + pk + b"\x03").digest()[:2] + b"\x03"
        clear = pk[:31] + bytes([pk[31] & 0x7F])
        assert (mine_p.match_key(clear) >= 0) == (theirs_p.match_key(clear) >= 0)
        assert (mine_p.match_address(ab) >= 0) == (theirs_p.match_address(ab) >= 0)
        assert (mine_g.match_address(ab) >= 0) == (theirs_g.match_address(ab) >= 0)


def test_bad_filters_raise():
    for bad in ["ab0", "suffix:abc", "suffix:abd", "[", "a]", "[e-b]", "???", "wat:abc", "",
                "regex:^(ab", "regex:^sh0p"]:
        with pytest.raises(_lib.OgError):
            filters.FilterSet.parse(bad)


def _address_text(pk):
    ab = key.address_bytes(pk)
    bits = "".join(format(b, "08b") for b in ab)
    return "".join(workload.ALPHABET[int(bits[5 * i:5 * i + 5], 2)] for i in range(56))


def _scan_cases(rng):
    """General filter sets planted on real addresses: every placement from 0 to
    the last, the sign symbol 49, the checksum symbols 51-55, '?' classes,
    several banks, and patterns longer than one bank (32 symbols)."""
    for i in range(400):
        pk = bytes(rng.getrandbits(8) for _ in range(32))
        text = _address_text(pk)
        kind = i % 5
        if kind == 0:
            L = rng.randint(3, 12)
            s = rng.randint(0, 56 - L)
            lines = ["contains:" + text[s:s + L]]
        elif kind == 1:
            L = rng.randint(3, 9)
            lines = ["suffix:" + text[56 - L:]]
        elif kind == 2:
            L = rng.randint(2, 6)
            s = rng.choice([44, 45, 46, 47, 48, 49, 50, 51])
            L = min(L, 56 - s)
            body = list(text[s:s + L])
            body[rng.randrange(L)] = "?"
            if all(c == "?" for c in body):
                body[0] = text[s]
            lines = ["contains:" + "".join(body)]
        elif kind == 3:
            L = rng.randint(33, 50)
            s = rng.randint(0, 56 - L)
            lines = ["contains:" + text[s:s + L], "prefix:" + text[:3] + "?" + text[4:40]]
        else:
            lines = ["contains:" + "".join(rng.choice(workload.ALPHABET) for _ in range(5))
                     for _ in range(12)]
            lines += ["suffix:" + text[52:], "contains:" + text[20:25]]
        yield pk, lines


def test_device_matcher_agrees_with_oracle_on_every_placement():
    """The automaton the kernels run (key pass, tail prefilter, address pass),
    executed on the host: its verdict equals the oracle's whole-address match;
    a key-pass hit is a hit whatever the sign; a hit the key pass did not
    settle is always flagged for the address stage."""
    rng = random.Random(2024)
    checked = hits = needed = 0
    for planted, lines in _scan_cases(rng):
        mine, theirs = filters.FilterSet.parse(lines), OracleFilters(lines)
        others = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(6)]
        for pk in [planted] + others:
            sign = pk[31] >> 7
            clear = pk[:31] + bytes([pk[31] & 0x7F])
            kh, need, hit = mine.scan(clear, sign)
            want = theirs.match_address(key.address_bytes(pk)) >= 0
            assert hit == want, (lines, pk.hex())
            assert mine.scan(clear, 1 - sign)[0] == kh
            if kh:
                assert hit
            if hit and not kh:
                assert need, (lines, pk.hex())
            checked += 1
            hits += hit
            needed += need
    assert hits >= 400 and needed > 0 and checked == 400 * 7


def test_tuning_cache_round_trip(tmp_path, monkeypatch):
    """The reference's file format: written atomically, read back, foreign
    entries kept, an unreadable file is no entry (src/tuning.rs:1-126)."""
    from onion_gen_amd import tuning
    p = tmp_path / "sub" / "tuning.yaml"
    monkeypatch.setenv("OG_TUNING_PATH", str(p))
    assert tuning.path() == str(p)
    assert tuning.get("device/x/hip/0/1.0.2") is None
    assert tuning.put("device/x/hip/0/1.0.2", 512, 262144, 2.9e10, 1700000000, 1, 10)
    assert tuning.put("device/y/opencl/0/1.0.2", 256, 65536, 1.5e9, 1700000001)
    got = tuning.get("device/x/hip/0/1.0.2")
    assert got["first"] == 512 and got["second"] == 262144 and got["rounds"] == 1
    assert abs(got["rate"] - 2.9e10) < 1e3 and got["at"] == 1700000000 and got["radix"] == 10
    text = p.read_text()
    assert text.startswith("format: 1") and "entries:" in text and "device/y/opencl/0/1.0.2" in text
    p.write_text("format: 2\nentries: {}\n")
    assert tuning.get("device/x/hip/0/1.0.2") is None


# The reference crate's version (Cargo.toml:3), which CARGO_PKG_VERSION puts at
# the end of every tuning key (src/tuning.rs:65-74).
REFERENCE_VERSION = "1.0.2"


def test_tuning_device_key_is_the_references():
    """og_tuning_device_key builds src/tuning.rs:65-74's key: the name with
    '/' and newline replaced, the api, the thread count and the version; the
    reference's own cases (src/tuning.rs:150-176) hold for it."""
    from onion_gen_amd import tuning
    assert tuning.device_key("AMD Instinct MI355X (gfx950:sramecc+:xnack-)", "hip", 0) == \
        "device/AMD Instinct MI355X (gfx950:sramecc+:xnack-)/hip/0/" + REFERENCE_VERSION
    assert tuning.device_key("a/b\nc", "cuda", 7) == "device/a_b_c/cuda/7/" + REFERENCE_VERSION
    assert tuning.device_key("a", "cuda", 1) != tuning.device_key("a", "cuda", 2)
    assert tuning.device_key("x", "metal", 1).endswith(REFERENCE_VERSION)
    forged, real = tuning.device_key("a/metal/1/9.9.9", "cuda", 1), tuning.device_key("a", "metal", 1)
    assert forged != real and "a/metal" not in forged
    # a host passes its own CARGO_PKG_VERSION
    assert tuning.set_version("2.0.0")
    try:
        assert tuning.device_key("x", "hip", 0) == "device/x/hip/0/2.0.0"
        assert not tuning.set_version("") and not tuning.set_version("1/2")
        assert tuning.device_key("x", "hip", 0).endswith("2.0.0")
    finally:
        assert tuning.set_version(REFERENCE_VERSION)


def test_tuning_choice_is_bounded_by_memory():
    """The walk's choice (og_tuning_choose): the fastest shape within the
    budget, then the smallest footprint within 3% of it.  The measured walk
    of a 6-symbol prefix on MI355X is the first case: 2^20 x 1024 chains
    (129 GB) at 3.37e10 against 2^19 x 1024 (64 GB) at 3.29e10 keys/s."""
    from onion_gen_amd import tuning
    GB = 1 << 30
    rates = [3.37e10, 3.29e10, 3.10e10, 2.50e10]
    sizes = [129 * GB, 64 * GB, 32 * GB, 8 * GB]
    assert tuning.choose(rates, sizes) == 1             # within 3%: the smaller
    assert tuning.choose(rates, sizes, tolerance=0.0) == 0
    assert tuning.choose(rates, sizes, tolerance=0.10) == 2
    assert tuning.choose(rates, sizes, budget=100 * GB) == 1
    assert tuning.choose(rates, sizes, budget=32 * GB) == 2
    assert tuning.choose(rates, sizes, budget=33 * GB, tolerance=0.25) == 3
    assert tuning.choose(rates, sizes, budget=4 * GB) is None
    assert tuning.choose([0.0, 0.0], [1, 2]) is None     # nothing measured
    # an unmeasured shape is never kept, whatever its size
    assert tuning.choose([3.0e10, 0.0], [10 * GB, 1 * GB]) == 0
    # equal footprints: the faster
    assert tuning.choose([3.0e10, 2.95e10], [GB, GB]) == 0


def test_host_checksum_matches_sha3():
    """The host Keccak of the address checksum (the device one is checked on
    the GPU) against hashlib's SHA3-256."""
    rng = random.Random(12)
    keys = [0, 2 ** 256 - 1] + [rng.getrandbits(256) for _ in range(500)]
    for k, g in zip(keys, host_field_op(20, keys, keys)):
        d = hashlib.sha3_256(b".onion checksum" + k.to_bytes(32, "little") + b"\x03").digest()
        assert g == d[0] | (d[1] << 8)


def test_key_files_in_tor_layout(tmp_path):
    """og_write_key (src/output.rs:42-88): three files with tor's headers, the
    secret readable by its owner only, an existing directory left alone."""
    import stat
    seed = bytes(range(32))
    secret = key.expanded_secret_at_offset(seed, 8 * 12345)
    pk = key.public_from_scalar(secret[:32])
    created, path = key.write_key(tmp_path / "keys", pk, secret)
    host = key.address(pk) + ".onion"
    assert created and path.endswith("/" + host)
    raw = open(os.path.join(path, "hs_ed25519_secret_key"), "rb").read()
    assert raw == b"== ed25519v1-secret: type0 ==\x00\x00\x00" + secret
    assert open(os.path.join(path, "hs_ed25519_public_key"), "rb").read() == \
        b"== ed25519v1-public: type0 ==\x00\x00\x00" + pk
    assert open(os.path.join(path, "hostname")).read() == host + "\n"
    assert stat.S_IMODE(os.stat(path).st_mode) == 0o700
    assert stat.S_IMODE(os.stat(os.path.join(path, "hs_ed25519_secret_key")).st_mode) == 0o600
    # the secret produces the public key (the check of src/verify.rs:120-125),
    # by textbook big-integer arithmetic
    from tests import ed25519_ref
    assert ed25519_ref.encode(ed25519_ref.mul(int.from_bytes(secret[:32], "little"))) == pk
    created2, path2 = key.write_key(tmp_path / "keys", pk, bytes(64))
    assert not created2 and path2 == path
    assert open(os.path.join(path, "hs_ed25519_secret_key"), "rb").read() == raw


def _dictionary_lines(rng, planted):
    """A dictionary of general forms too large for the automaton banks:
    substrings (some with '?' and classes), suffixes and class prefixes, with
    some taken from the planted keys' addresses so that every kind hits."""
    A = workload.ALPHABET
    lines = []
    for _ in range(700):
        w = [rng.choice(A) for _ in range(rng.randint(4, 6))]
        if rng.random() < 0.1:
            w[rng.randrange(len(w))] = "?"
        if rng.random() < 0.1:
            w[rng.randrange(len(w))] = "[a-e]"
        lines.append("contains:" + "".join(w))
    for _ in range(200):
        lines.append("suffix:" + "".join(rng.choice(A) for _ in range(rng.randint(2, 6))) +
                     rng.choice("aiqy") + "d")
    for _ in range(40):
        lines.append("[a-d]" + "".join(rng.choice(A) for _ in range(4)))
    for pk in planted:
        text = _address_text(pk)
        kind = rng.randrange(3)
        if kind == 0:
            s = rng.randint(0, 50)
            lines.append("contains:" + text[s:s + 6])
        elif kind == 1:
            lines.append("suffix:" + text[56 - rng.randint(3, 8):])
        else:
            lines.append("[" + text[0] + "z]" + text[1:5])
    return lines


def _check_dictionary(lines, keys):
    mine, theirs = filters.FilterSet.parse(lines), OracleFilters(lines)
    assert any(l.startswith("dictionary automat") for l in mine.describe())
    mine_prefiltered = any(l.startswith("window prefilter") for l in mine.describe())
    hits = needed = key_hits = 0
    prev = None
    for pk in keys:
        sign = pk[31] >> 7
        clear = pk[:31] + bytes([pk[31] & 0x7F])
        kh, need, hit = mine.scan(clear, sign)
        want = theirs.match_address(key.address_bytes(pk)) >= 0
        if prev is not None:
            # the kernels' pair form: the single scan's verdicts, except that
            # behind the window prefilter a key may go to the address stage in
            # place of a key-stage verdict - never the other way round - and
            # the address stage then settles it
            pair = mine.scan_pair(prev[0], clear)
            for (skh, sneed, swant), (gkh, gneed) in zip((prev[1], (kh, need, want)), pair):
                if not mine_prefiltered:
                    assert (gkh, gneed) == (skh, sneed)
                assert not gkh or skh
                assert not (skh or sneed) or gkh or gneed
                assert (gkh or (gneed and swant)) == swant
        prev = (clear, (kh, need, want))
        assert hit == want, pk.hex()
        if kh:
            assert hit and mine.scan(clear, 1 - sign)[0]
        if hit and not kh:
            assert need, pk.hex()
        hits += hit
        needed += need
        key_hits += kh
    return hits, needed, key_hits


def test_dictionary_automaton_agrees_with_oracle():
    """General forms beyond the banks run as one automaton over the address
    (the reference's flat substring automaton and tail prefilter,
    src/filter.rs:1044-1207): its verdict equals the oracle's on every key, a
    key-stage hit is a hit whatever the sign, and an unsettled hit is always
    sent to the address stage."""
    rng = random.Random(77)
    planted = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(60)]
    keys = planted + [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(1000)]
    hits, needed, key_hits = _check_dictionary(_dictionary_lines(rng, planted), keys)
    assert hits >= 60 and key_hits > 0


def test_dictionary_tail_liveness_is_selective():
    """Eight-symbol substrings can reach the checksum only from symbol 48 on,
    where the key, the sign and the fixed tail rule nearly every placement out:
    the automaton's tail liveness sends almost no key to the checksum.  (The
    reference's tail prefilter passes every candidate once a needle of seven
    symbols or fewer exists, src/filter.rs:1058-1085; here symbols 49-51 and the
    a/i/q/y, d tail still count.)"""
    rng = random.Random(78)
    A = workload.ALPHABET
    planted = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(40)]
    lines = ["contains:" + "".join(rng.choice(A) for _ in range(8)) for _ in range(1500)]
    for pk in planted:
        s = rng.randint(0, 41)
        lines.append("contains:" + _address_text(pk)[s:s + 8])
    keys = planted + [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(800)]
    hits, needed, key_hits = _check_dictionary(lines, keys)
    assert hits >= 40 and key_hits >= 40
    assert needed < 0.05 * len(keys), needed


def test_dictionary_beyond_one_automaton_splits():
    """10,000 eight-symbol substrings need about 50,000 automaton states, past
    the 32,767 one automaton holds: the set is split into several automata,
    and the verdicts still equal the oracle's."""
    rng = random.Random(79)
    A = workload.ALPHABET
    planted = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(20)]
    lines = ["contains:" + "".join(rng.choice(A) for _ in range(8)) for _ in range(10000)]
    for pk in planted:
        s = rng.randint(0, 41)
        lines.append("contains:" + _address_text(pk)[s:s + 8])
    mine = filters.FilterSet.parse(lines)
    assert any(l.startswith("dictionary automata: 2 of") for l in mine.describe())
    keys = planted + [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(100)]
    hits, needed, key_hits = _check_dictionary(lines, keys)
    assert hits >= 20 and key_hits >= 20


def test_window_prefilter_and_the_address_stage_give_the_whole_walks_result():
    """A dictionary of plain substrings (1,000 eight-symbol words, one
    automaton in global memory) runs the key stage behind the window
    prefilter: a short walk over symbols 42..48 for every key, and a key whose
    7-symbol window hits both maps goes to the address stage, which decides on
    the whole address.  On random keys and on keys with a word planted at
    every placement, across the key/checksum boundary included:
    - a key-stage hit is a real one (the whole walk's too);
    - the whole walk's hits and open keys always reach the address stage or
      are hits, so the key stage then the address stage report exactly the
      keys whose address holds a word;
    - random keys go to the address stage for a window hit about as often as
      the maps' occupancy predicts, not for every key."""
    import random
    from onion_gen_amd import filters, key
    rng = random.Random(31)
    alphabet = "abcdefghijklmnopqrstuvwxyz234567"
    words = ["".join(rng.choice(alphabet) for _ in range(8)) for _ in range(1000)]
    fs = filters.FilterSet.parse(["contains:" + w for w in words])
    assert any("window prefilter" in line for line in fs.describe())
    # a set that cannot take it (a class in a word) keeps the plain walk
    fs2 = filters.FilterSet.parse(["contains:" + w for w in words[:999]] + ["contains:ab[cd]efghi"])
    assert not any("window prefilter" in line for line in fs2.describe())
    wordset = set(words)

    def key_with(word, at):
        bits = [rng.getrandbits(1) for _ in range(255)]
        for i, c in enumerate(word):
            s = at + i
            if s >= 51:  # symbols 49 and 50 are key bits too (49 holds the sign)
                break
            v = alphabet.index(c)
            for b in range(5):
                if 5 * s + b < 255:
                    bits[5 * s + b] = (v >> (4 - b)) & 1
        # the address reads the packed key's bytes in order, each from its top
        # bit; bit 248 of that string (the top bit of byte 31) is the sign
        pk = bytearray((int("".join(map(str, bits)), 2) << 1).to_bytes(32, "big"))
        pk[31] &= 0x7F
        return bytes(pk)

    def holds_word(pk):
        addr = key.address(pk)[:56]
        return any(addr[s:s + 8] in wordset for s in range(49))

    random_keys = [bytes(rng.getrandbits(8) for _ in range(31)) + bytes([rng.getrandbits(7)])
                   for _ in range(2000)]
    keys = list(random_keys)
    for at in range(0, 51):
        for _ in range(8):
            keys.append(key_with(rng.choice(words), at))
    checked = hits = 0
    sent = 0
    for a, b in zip(keys[0::2], keys[1::2]):
        pair = fs.scan_pair(a, b)
        for i, k in enumerate((a, b)):
            kh, nd = pair[i]
            kh1, nd1, _ = fs.scan(k, 0)
            assert not kh or kh1, k.hex()              # a key-stage hit is real
            assert not (kh1 or nd1) or kh or nd, k.hex()  # never closed where it matters
            assert (kh or (nd and holds_word(k))) == holds_word(k), k.hex()
            checked += 1
            hits += holds_word(k)
            sent += nd and not (kh1 or nd1) and k in random_keys
    assert checked == len(keys) - len(keys) % 2 and hits > 150
    # 1,000 words: about 5.8e-5 per probe, 22 probes, so about 0.13% of keys
    assert sent <= 20, sent


def _address_with(rng, word, at):
    """35 random address bytes with `word` at symbol `at` of the 56."""
    alphabet = workload.ALPHABET
    bits = [rng.getrandbits(1) for _ in range(280)]
    for i, c in enumerate(word):
        v = alphabet.index(c)
        for b in range(5):
            bits[5 * (at + i) + b] = (v >> (4 - b)) & 1
    return int("".join(map(str, bits)), 2).to_bytes(35, "big")


def test_address_window_prefilter_never_stops_a_matching_walk():
    """1,000 five-symbol substrings send nearly every candidate to the address
    stage, which then probes every 5-symbol window of the 56 symbols against
    two maps and walks the automaton only where one hits.  On addresses with a
    word planted at every placement (the checksum and version bytes included)
    the walk always runs and matches; on random addresses it runs about as
    rarely as the maps' occupancy predicts."""
    rng = random.Random(41)
    words = sorted({"".join(rng.choice(workload.ALPHABET) for _ in range(5)) for _ in range(1000)})
    fs = filters.FilterSet.parse(["contains:" + w for w in words])
    lines = fs.describe()
    assert any(l.startswith("window prefilter in the address stage") for l in lines), lines
    assert any(l.startswith("key stage skipped") for l in lines), lines
    wordset = set(words)
    for at in range(52):
        for _ in range(6):
            addr = _address_with(rng, rng.choice(words), at)
            window, match = fs.address_window(addr)
            assert match and window, (at, addr.hex())
    walks = matches = 0
    for _ in range(4000):
        addr = bytes(rng.getrandbits(8) for _ in range(35))
        window, match = fs.address_window(addr)
        text = "".join(workload.ALPHABET[(int.from_bytes(addr, "big") >> (275 - 5 * j)) & 31]
                       for j in range(56))
        assert match == any(text[s:s + 5] in wordset for s in range(52))
        assert window or not match
        walks += window
        matches += match
    # about 0.23% of random addresses pass: 52 probes of about 4.4e-5 each
    assert walks <= 40 and walks >= matches, (walks, matches)


def _key_with_prefix(word, rng):
    """A canonical packed key (sign clear) whose address starts with `word`."""
    bits = 5 * len(word)
    v = int("".join(format(workload.ALPHABET.index(c), "05b") for c in word) or "0", 2)
    k = (v << (256 - bits)) | (rng.getrandbits(256) & ((1 << (256 - bits)) - 1))
    pk = bytearray(k.to_bytes(32, "big"))
    pk[31] &= 0x7F
    return bytes(pk)


def test_exact_table_matches_the_sets_own_matcher():
    """The device's exact table behind the bitmap (og_scan.hpp exact_table:
    a directory over each length's sorted keys, four lengths side by side),
    run on the host, against FilterSet::match_key (src/filter.rs:2194-2217):
    on keys that start with a word of the set, with a word whose last symbol
    is changed, and at random.  A hit of the set's matcher is always a hit of
    the table; for words of up to 12 symbols the two agree, and a longer word
    is compared on its first 64 bits (a "maybe")."""
    rng = random.Random(21)
    mixed = set()
    while len(mixed) < 400:
        n = rng.choice([2, 3, 4, 5, 6, 7, 9, 11, 12, 13, 14])
        mixed.add("".join(rng.choice(workload.ALPHABET) for _ in range(n)))
    for words in (workload.dictionary(10000, seed=2), sorted(mixed)):
        fs = filters.FilterSet.parse(words)
        # a long word's first 64 bits: its first 12 symbols and the top 4 bits
        # of the 13th
        heads = {(w[:12], workload.ALPHABET.index(w[12]) >> 1) for w in words if len(w) >= 13}
        keys = []
        for w in rng.sample(words, min(600, len(words))):
            keys.append(_key_with_prefix(w, rng))
            other = rng.choice([c for c in workload.ALPHABET if c != w[-1]])
            keys.append(_key_with_prefix(w[:-1] + other, rng))
        keys += [_key_with_prefix("", rng) for _ in range(3000)]
        hits = 0
        for pk in keys:
            table, match = fs.exact_table(pk), fs.match_key(pk) >= 0
            hits += match
            assert table or not match, pk.hex()
            if table and not match:
                head = int.from_bytes(pk[:8], "big")
                sym = "".join(workload.ALPHABET[(head >> (59 - 5 * i)) & 31] for i in range(12))
                assert (sym, head & 0xF) in heads, pk.hex()
        assert hits >= min(600, len(words))


@pytest.mark.parametrize("count", [1, 2])
def test_literal_bank_agrees_with_oracle_at_every_placement(count):
    """A bank of one or two literals of 1..6 symbols runs the key pass as a
    compare at each placement (og_scan.hpp literal_min_from) and the tail from
    the last six symbols' automaton: on keys with a literal planted at every
    placement 0..55, with its last symbol changed, and at random, the host run
    of the kernels' matcher equals the oracle's whole-address match, and a hit
    the key pass leaves unsettled is flagged for the address stage."""
    rng = random.Random(77 + count)
    for L in range(1, 7):
        for trial in range(3):
            pk0 = bytes(rng.getrandbits(8) for _ in range(32))
            text = _address_text(pk0)
            lit = "".join(rng.choice(workload.ALPHABET) for _ in range(L))
            lines = ["contains:" + lit]
            if count == 2:
                other = "".join(rng.choice(workload.ALPHABET) for _ in range(rng.randint(2, 6)))
                lines.append("contains:" + other)
            mine, theirs = filters.FilterSet.parse(lines), OracleFilters(lines)
            assert any(l.startswith("literal compare: 1") for l in mine.describe()), lines
            keys = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(40)]
            for s in range(0, 56 - L + 1):
                for word in (lit, lit[:-1] + rng.choice([c for c in workload.ALPHABET if c != lit[-1]])):
                    if s + L > 51:
                        break  # the checksum symbols are not set by the key
                    # symbols 0..50 are the key's bits 0..254 (big-endian);
                    # bit 255 stays the original key's
                    t = text[:s] + word + text[s + L:]
                    bits = "".join(format(workload.ALPHABET.index(c), "05b") for c in t[:51])
                    keys.append(int(bits + str(pk0[31] & 1), 2).to_bytes(32, "big"))
            hits = 0
            for pk in keys:
                sign = pk[31] >> 7
                clear = pk[:31] + bytes([pk[31] & 0x7F])
                kh, need, hit = mine.scan(clear, sign)
                want = theirs.match_address(key.address_bytes(pk)) >= 0
                assert hit == want, (lines, pk.hex())
                if kh:
                    assert hit
                if hit and not kh:
                    assert need, (lines, pk.hex())
                hits += hit
            assert hits >= 56 - L - 5
