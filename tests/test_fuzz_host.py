"""Fuzzing the host parsers on the CPU: the filter syntax, the `regex:` form
and the tuning file read user input, so arbitrary input must end in a set or
an error message, never a crash.  Under the sanitizer build
(`make -C onion_gen_amd asan-test`) an out-of-bounds access or undefined
behaviour anywhere on these paths fails the run.  Seeded, so a failure
reproduces.
"""
import os
import random

from onion_gen_amd import filters, key, tuning
from onion_gen_amd._lib import OgError

REGEX_ATOMS = list("abcdxyz234567") + ["^", "$", ".", "*", "+", "?", "|", "(", ")", "[", "]",
                                       "{", "}", ",", "-", "\\", "\\b", "\\B", "\\d", "\\w",
                                       "\\z", "\\A", "(?i)", "(?x)", "(?u)", "[[:digit:]]",
                                       "&&", "--", "~~", "[^", "a{2,", "{3}", "{,9}", " ",
                                       "#", "\n", "\x00"[:0], "é", "\\x41", "\\p{L}"]
FILTER_ATOMS = list("abcdefgz2567?") + ["[", "]", "-", "[a-d]", "[^", "prefix:", "suffix:",
                                        "contains:", "regex:", ":", " ", "D", "Z", "0", "1",
                                        "89", "\t"]


def rand_text(rng, atoms, lo, hi):
    return "".join(rng.choice(atoms) for _ in range(rng.randint(lo, hi)))


def random_address(rng):
    return bytes(rng.getrandbits(8) for _ in range(32))


def try_parse(lines):
    try:
        return filters.FilterSet.parse(lines)
    except OgError as e:
        assert str(e)  # every refusal says why
        return None


def exercise(fs, rng):
    fs.describe()
    for _ in range(4):
        pk = random_address(rng)
        ab = key.address_bytes(pk)
        fs.match_address(ab)
        fs.match_key(pk[:31] + bytes([pk[31] & 0x7F]))
        try:
            fs.scan(pk[:31] + bytes([pk[31] & 0x7F]), pk[31] >> 7)
        except OgError:
            pass  # a set without general forms has no device matcher to run


def test_fuzz_regex_form():
    rng = random.Random(11)
    parsed = 0
    for _ in range(2500):
        fs = try_parse(["regex:" + rand_text(rng, REGEX_ATOMS, 0, 24)])
        if fs is not None:
            parsed += 1
            exercise(fs, rng)
    assert parsed > 100


def test_fuzz_filter_lines():
    rng = random.Random(12)
    parsed = 0
    for _ in range(2500):
        lines = [rand_text(rng, FILTER_ATOMS, 0, 12) for _ in range(rng.randint(1, 4))]
        fs = try_parse(lines)
        if fs is not None:
            parsed += 1
            exercise(fs, rng)
    assert parsed > 40


def test_fuzz_long_and_odd_inputs():
    rng = random.Random(13)
    cases = ["", "\n\n", "regex:", "regex:" + "(" * 500, "regex:" + "a{1000}" * 3,
             "regex:" + "[" * 200, "x" * 10000, "contains:" + "a" * 60,
             "suffix:" + "a" * 57 + "d", "regex:(a|b)*" * 50, "[" + "a-z" * 100 + "]",
             "regex:\\", "regex:(?", "regex:a{99999999999}", "regex:[z-a]"]
    cases += ["regex:" + "".join(chr(rng.randint(1, 255)) for _ in range(30)) for _ in range(200)]
    for text in cases:
        fs = try_parse([text])
        if fs is not None:
            exercise(fs, rng)


def test_fuzz_tuning_file(tmp_path, monkeypatch):
    """Truncated, foreign and random cache files read as no entry or as an
    entry; the next put replaces what it cannot read (src/tuning.rs:1-18)."""
    rng = random.Random(14)
    path = tmp_path / "tuning.yaml"
    monkeypatch.setenv("OG_TUNING_PATH", str(path))
    good = ("format: 1\nentries:\n  'device/x/hip/0/1.0.2':\n    first: 512\n    second: 262144\n"
            "    rate: 3e10\n    at: 1\n    slots: 64\n")
    pieces = good.splitlines(keepends=True) + ["  :\n", "    :\n", "format: 1", "'", "\t", "\x00",
                                               "    first: -5\n", "    rate: nan\n",
                                               "  " + "k" * 5000 + ":\n", "entries:\n"]
    for i in range(400):
        if i % 3 == 0:
            body = good[:rng.randint(0, len(good))]
        else:
            body = "".join(rng.choice(pieces) for _ in range(rng.randint(0, 12)))
        path.write_bytes(body.encode("utf-8", "surrogatepass"))
        got = tuning.get("device/x/hip/0/1.0.2")
        assert got is None or isinstance(got["first"], int)
    os.remove(path)
    assert tuning.put("device/x/hip/0/1.0.2", 512, 262144, 3e10, 1, 1, 10, 64)
    assert tuning.get("device/x/hip/0/1.0.2")["slots"] == 64
