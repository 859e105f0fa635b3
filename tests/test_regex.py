"""CPU tests of the regex: form (src/expression.rs, src/filter.rs:353-375, 884-1005).

The reference matches with the `regex` crate (1.13.1) and extracts its guard
literals with `regex-syntax` (0.8.11); neither is importable here.  The ground
truth is Python's `re`, an independent engine that reads the same syntax for
every expression below (the reference's own corpus from
tests/regex_prefilter_is_sound.rs:22-62, plus more), applied to the finished
address with nothing of ours in between.  Syntax only the crate has (\\z, \\<,
[[:digit:]], class set operations) is compared with a Python expression that
means the same over base32 text.
"""
import random
import re
import zlib

import pytest

from onion_gen_amd import _lib, filters, key, workload

# tests/regex_prefilter_is_sound.rs:22-62, then the boundary cases of looks
# (\b and \B at both ends), flags and the reference's run tests
# (tests/regex_matching.rs:95-120).
CORPUS = [
    "^shop", "^shop[a-z]{3}", "^shop.*", "^(shop|store)", "^my(shop|store)",
    "^(a|b|c)[a-z]{3}", "^s[a-z]op", "^sh.p", "^shop.*store", "^[a-z]{4}shop", "^zz+",
    "^(shop|shopping)", "^[bcdfghjklmnpqrstvwxyz]{6}", "^a[2-7]{3}shop", "^(ab)?cd", "^x*yz",
    "^(q|)rs", "^[a-z]{2}[2-7]{2}[a-z]{2}", "shop", "(cat|dog)food", "sh.p", "[2-7]{4}", "z{3}",
    "shop$", "^.*shop$", "[a-z]d$", "^s.*p$", "(?i)ABC", "^(a|ab|abc)x", "^..shop",
    "^[a-z]shop[2-7]", "qqq|www", "^(?:a[2-7]|b[2-7])c", "^\\w{4}", "^a.{10}z",
    "^[a-z]{49}$", "^[a-z]{49}\\b", "^.{49}$", "^[a-z]{48}[2-7]$", "^\\w{49}\\b", "^[a-z]{20}$",
    "a\\Bb", "b\\B", "\\Ba", "x\\b", "^zz.*qd$", "zz[2-7]z", "^ab[2-7]", "^[bcdfghjklmnp]{4}",
    "[a-z]{12}", "(ab|cd)*ef", "^(?i:AB)c", "[^a-y2-6]{3}", "^\\d\\d", "\\D{50}", "a{2,3}?b",
    "^[a-c]{,2}d", "qd$", "[aeiou]{3}[2-7]", "(?s).", "^(abc|0xy)", "ab|^cd", "(ab)+$",
]

# Crate syntax Python lacks, with a Python expression that means the same over
# base32 text (all of whose symbols are word characters).
TRANSLATED = [
    ("\\Ashop", "^shop"), ("shop\\z", "shop$"), ("[[:digit:]]{3}", "[0-9]{3}"),
    ("[a-z&&[^aeiou]]{4}", "[b-df-hj-np-tv-z]{4}"), ("[a-z--[b-y]]{2}", "[az]{2}"),
    ("[a-f~~c-h]{2}", "[abgh]{2}"), ("\\<ab", "^ab"), ("qd\\>", "qd$"),
    ("\\b{start}ab", "^ab"), ("ad\\b{end-half}", "ad$"), ("(?<word>ab)c", "abc"),
    ("\\x61b", "ab"), ("\\x{62}c", "bc"), ("(?U)ab+", "ab+"), ("(?x) a b  c # spaced", "abc"),
    ("[[:^alpha:]]{2}", "[0-9]{2}"), ("(?i)[^A-X]{3}", "[yz2-7]{3}"),
]

SEEDS = ["shop", "store", "catfood", "abc", "zzz", "myshop", "zz3z", "qd", "ab3", "bcdf", "ef",
         "cdef", "ab", "cd", "yz"]


def _text_to_bytes(text):
    """56 base32 symbols are exactly the 35 address bytes."""
    v = 0
    for ch in text:
        v = (v << 5) | workload.ALPHABET.index(ch)
    return v.to_bytes(35, "big")


def _random_texts(n, seed):
    """Random base32 text, not protocol-shaped, so $-anchored expressions see
    varied tails (tests/regex_prefilter_is_sound.rs:183-205), with seeded
    words planted often enough that every expression meets positives."""
    rng = random.Random(seed)
    out = []
    for i in range(n):
        s = [rng.choice(workload.ALPHABET) for _ in range(56)]
        if i % 3 == 0:
            w = rng.choice(SEEDS)
            at = 0 if i % 2 == 0 else rng.randrange(0, 57 - len(w))
            s[at:at + len(w)] = list(w)
        if i % 7 == 0:
            s[-2:] = ["q", "d"]
        out.append("".join(s))
    return out


def _address_text(pk):
    ab = key.address_bytes(pk)
    bits = "".join(format(b, "08b") for b in ab)
    return "".join(workload.ALPHABET[int(bits[5 * i:5 * i + 5], 2)] for i in range(56))


def _set_symbol(packed, i, value):
    """Writes symbol i (< 49, where the key settles it) into the packed key
    (tests/regex_prefilter_is_sound.rs:318-338)."""
    bit = i * 5
    byte, shift = bit // 8, bit % 8
    window = (packed[byte] << 8) | (packed[byte + 1] if byte + 1 < 32 else 0)
    pos = 11 - shift
    window = (window & ~(0x1F << pos)) | ((value & 0x1F) << pos)
    packed[byte] = (window >> 8) & 0xFF
    if byte + 1 < 32:
        packed[byte + 1] = window & 0xFF


def _random_keys(n, seed):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        pk = bytearray(rng.getrandbits(8) for _ in range(32))
        if i % 4 == 0:
            w = rng.choice(SEEDS)
            at = 0 if i % 8 == 0 else rng.randrange(0, 49 - len(w))
            for j, ch in enumerate(w):
                _set_symbol(pk, at + j, workload.ALPHABET.index(ch))
        out.append(bytes(pk))
    return out


def test_planted_symbols_read_back():
    pk = bytearray([0x5A] * 32)
    for at in (0, 1, 3, 7, 8, 12, 47, 48):
        for v in range(32):
            _set_symbol(pk, at, v)
            assert _address_text(bytes(pk))[at] == workload.ALPHABET[v]


def test_host_matcher_agrees_with_python_re():
    texts = _random_texts(4000, 11)
    with_positives = 0
    for source in CORPUS:
        fs = filters.FilterSet.parse("regex:" + source)
        truth = re.compile(source)
        positives = 0
        for text in texts:
            want = truth.search(text) is not None
            got = fs.match_address(_text_to_bytes(text)) >= 0
            assert got == want, (source, text)
            positives += want
        with_positives += positives > 0
    # a comparison where nothing matched would prove nothing
    assert with_positives >= len(CORPUS) * 3 // 4


@pytest.mark.parametrize("source,python", TRANSLATED)
def test_crate_syntax_agrees_with_its_python_equivalent(source, python):
    fs = filters.FilterSet.parse("regex:" + source)
    truth = re.compile(python)
    positives = 0
    for text in _random_texts(1500, 99):
        want = truth.search(text) is not None
        assert (fs.match_address(_text_to_bytes(text)) >= 0) == want, (source, text)
        positives += want
    assert positives > 0, source


@pytest.mark.parametrize("source", CORPUS + [s for s, _ in TRANSLATED])
def test_device_matcher_agrees_with_the_expression_applied_directly(source):
    """The proxies, the automata and the key stage the kernels run, executed on
    the host (tests/regex_prefilter_is_sound.rs:208-280): the verdict equals the
    expression on the finished address; a key-stage hit is a hit whatever the
    sign; a match the key stage did not settle is always sent to the address
    stage."""
    python = dict(TRANSLATED).get(source, source)
    fs = filters.FilterSet.parse("regex:" + source)
    truth = re.compile(python)
    pairs = not any(l.startswith("key tables") for l in fs.describe())
    rng = random.Random(5)
    prev = None
    for pk in _random_keys(1200, zlib.crc32(source.encode())):
        sign = rng.getrandbits(1)
        full = pk[:31] + bytes([(pk[31] & 0x7F) | (sign << 7)])
        clear = pk[:31] + bytes([pk[31] & 0x7F])
        want = truth.search(_address_text(full)) is not None
        kh, need, hit = fs.scan(clear, sign)
        if pairs and prev is not None:
            # the kernels run a pair's two candidates through one walk
            assert fs.scan_pair(prev[0], clear) == (prev[1], (kh, need)), source
        prev = (clear, (kh, need))
        assert hit == want, (source, full.hex())
        if kh:
            assert want and fs.scan(clear, 1 - sign)[0]
        if want and not kh:
            assert need, (source, full.hex())


def test_a_mixed_set_reports_both_kinds():
    """tests/regex_matching.rs:136: a prefix beside an expression; the
    expression comes after the other filters (src/filter.rs:1920)."""
    fs = filters.FilterSet.parse(["zz", "regex:^ab[2-7]"])
    assert len(fs) == 2
    assert fs.match_address(_text_to_bytes("zz" + "a" * 54)) == 0
    assert fs.match_address(_text_to_bytes("ab4" + "a" * 53)) == 1
    assert fs.match_address(_text_to_bytes("ab9"[:2] + "b" * 54)) == -1


def test_expressions_are_never_absorbed():
    # src/filter.rs:869-876: `^a` must not swallow `^ab|cd`
    fs = filters.FilterSet.parse(["regex:^a", "regex:^ab|cd", "a"])
    assert len(fs) == 3
    assert fs.match_address(_text_to_bytes("cd" + "b" * 54)) == 2


def test_proxies_are_the_obligatory_sequences():
    """src/expression.rs:411-470: what the extraction proves, here as symbol
    sets (one sequence where the reference lists 216 literals)."""
    def proxies(source):
        line = [l for l in filters.FilterSet.parse("regex:" + source).describe()
                if l.startswith("regex:")][0]
        return line
    assert "^shop," in proxies("^shop") and "first 49 symbols" in proxies("^shop")
    assert "^myshop | ^mystore" in proxies("^my(shop|store)")
    assert "catfood | dogfood" in proxies("(cat|dog)food")
    assert "SHA3" in proxies("(cat|dog)food")
    assert "^a[234567][234567][234567]shop" in proxies("^a[2-7]{3}shop")
    assert "^shop," in proxies("^shop.*") and "SHA3" in proxies("^shop.*")
    assert "no proxy" in proxies("[a-z]d$")
    assert "^[bcdfghjklmnp][bcdfghjklmnp][bcdfghjklmnp][bcdfghjklmnp]" in proxies(
        "^[bcdfghjklmnp]{4}")
    # `x*` lets the expression start at `y`: both branches are kept
    assert "^x | ^yz" in proxies("^x*yz")
    assert "SHA3" in proxies("^[a-z]{20}$") and "SHA3" in proxies("^[a-z]{49}\\b")


@pytest.mark.parametrize("source,needle", [
    ("^sh0p", "'0'"),                              # src/expression.rs:530-540
    ("^[a-z]{57}", "57"),                          # :550-560
    ("[^\\x00-\\xff]", "no address can ever match"),  # :520-527
    ("^(shop", "does not parse"),                  # :562-571
    ("(?=ab)", "look-around"),
    ("(a)\\1", "backreferences"),
    ("\\pL", "Unicode"),
    ("a**", "repetition operator missing expression"),
    ("[b-a]", "invalid character class range"),
    ("a{3,2}", "invalid repetition count range"),
])
def test_unmatchable_or_malformed_expressions_are_refused(source, needle):
    with pytest.raises(_lib.OgError, match=re.escape(needle)):
        filters.FilterSet.parse("regex:" + source)
    if source == "^[a-z]{57}":
        with pytest.raises(_lib.OgError, match="56"):
            filters.FilterSet.parse("regex:" + source)


def test_one_live_branch_keeps_the_expression():
    # src/expression.rs:541-547
    fs = filters.FilterSet.parse("regex:^(abc|0xy)")
    assert fs.match_address(_text_to_bytes("abc" + "a" * 53)) == 0


def test_form_name_is_case_insensitive_and_body_trimmed():
    # src/filter.rs:489-492 strip_form
    fs = filters.FilterSet.parse("  REGEX :  ^ab  ")
    assert fs.match_address(_text_to_bytes("ab" + "c" * 54)) == 0
    with pytest.raises(_lib.OgError, match="regex"):
        filters.FilterSet.parse("pattern:abc")
