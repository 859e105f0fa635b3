"""Workload conventions: the seeded inputs the parity tests and the bench use.

"ChaCha20 seeded n" follows rand's ``ChaCha20Rng::seed_from_u64(n)``: rand_core's
PCG32 (XSH RR) expansion of n into a 32-byte key, then the ChaCha20 keystream
(RFC 8439 block function, nonce 0, block counter from 0).
"""
import struct

ALPHABET = "abcdefghijklmnopqrstuvwxyz234567"
_MASK = 0xFFFFFFFF


def _pcg32_key(seed):
    state = seed & 0xFFFFFFFFFFFFFFFF
    out = b""
    for _ in range(8):
        state = (state * 6364136223846793005 + 11634580027462260723) & 0xFFFFFFFFFFFFFFFF
        xs = (((state >> 18) ^ state) >> 27) & _MASK
        rot = state >> 59
        x = ((xs >> rot) | (xs << ((32 - rot) & 31))) & _MASK
        out += struct.pack("<I", x)
    return out


def _rotl(x, n):
    return ((x << n) | (x >> (32 - n))) & _MASK


def _block(key_words, counter):
    s = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(key_words) + [counter, 0, 0, 0]
    x = list(s)

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]) & _MASK; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & _MASK; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & _MASK; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & _MASK; x[b] = _rotl(x[b] ^ x[c], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return struct.pack("<16I", *[(x[i] + s[i]) & _MASK for i in range(16)])


def chacha_stream(seed, n):
    """The first ``n`` keystream bytes of ChaCha20 seeded ``seed``."""
    key = struct.unpack("<8I", _pcg32_key(seed))
    out = bytearray()
    counter = 0
    while len(out) < n:
        out += _block(key, counter)
        counter += 1
    return bytes(out[:n])


def symbols(stream, count, offset=0):
    return "".join(ALPHABET[b & 31] for b in stream[offset:offset + count])


# ---- the configurations (their sizes are the bench's and the parity tests')

def config1():
    """Plain CPU case: 4-symbol prefix (seed 1), secret seed from ChaCha20(0),
    Montgomery width 256, 2^24 sequential candidates."""
    return {"seed": chacha_stream(0, 32), "prefix": symbols(chacha_stream(1, 8), 4),
            "width": 256, "candidates": 1 << 24}


def config2():
    """Same method on the GPU: one 6-symbol prefix (seed 1), 2^34 candidates."""
    return {"seed": chacha_stream(0, 32), "prefix": symbols(chacha_stream(1, 8), 6),
            "candidates": 1 << 34}


def dictionary(count=10000, seed=2):
    """``count`` distinct prefixes of length uniform in 5..8 from ChaCha20(seed):
    per filter one byte picks the length (5 + b % 4), then one byte per symbol."""
    stream = chacha_stream(seed, count * 40)
    seen, out, i = set(), [], 0
    while len(out) < count:
        if i + 9 > len(stream):
            stream += chacha_stream(seed, len(stream) * 2)[len(stream):]
        length = 5 + stream[i] % 4
        word = symbols(stream, length, i + 1)
        i += 1 + length
        if word not in seen:
            seen.add(word)
            out.append(word)
    return out


# ---- the batched and served mixes

def _suffix(stream, i, length):
    body = [ALPHABET[b & 31] for b in stream[i:i + length]]
    body[-1] = "d"                              # the version byte fixes the last symbol
    body[-2] = "aiqy"[stream[i + length] & 3]   # and leaves four values before it
    return "".join(body)


def search_mix(count, seed, hits="loguniform"):
    """``count`` searches: (seed32, filter text, wanted hits).  Filter kind is
    uniform over {prefix of 4-7 symbols, suffix of 3-5 symbols, substring of 5};
    wanted hits are log-uniform in 1..1000 (``hits="loguniform"``) or uniform
    over {1, 10, 100} (``hits="choice"``)."""
    import math
    stream = chacha_stream(seed, count * 64)
    out = []
    for r in range(count):
        s = stream[64 * r:64 * r + 64]
        kind = s[0] % 3
        if kind == 0:
            text = symbols(s, 4 + s[1] % 4, 8)
        elif kind == 1:
            text = "suffix:" + _suffix(s, 8, 3 + s[1] % 3)
        else:
            text = "contains:" + symbols(s, 5, 8)
        if hits == "loguniform":
            u = int.from_bytes(s[2:6], "little") / 2 ** 32
            want = max(1, min(1000, int(round(math.exp(u * math.log(1000))))))
        else:
            want = (1, 10, 100)[s[2] % 3]
        out.append((s[32:64], text, want))
    return out


def poisson_arrivals(rate, duration, seed):
    """Arrival times of a Poisson process of ``rate`` per second over
    ``duration`` seconds, from ChaCha20(seed) uniforms."""
    import math
    times, t, i = [], 0.0, 0
    stream = chacha_stream(seed, int(rate * duration * 3 + 64) * 8)
    while True:
        if 8 * i + 8 > len(stream):
            stream += chacha_stream(seed + 1_000_003 * (i + 1), len(stream))
        u = (int.from_bytes(stream[8 * i:8 * i + 8], "little") + 0.5) / 2 ** 64
        i += 1
        t += -math.log(u) / rate
        if t >= duration:
            return times
        times.append(t)
