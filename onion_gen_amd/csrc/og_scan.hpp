// The general filter forms (suffix, substring, class prefix) as one
// bit-parallel Shift-And automaton over the 56 address symbols, for the host
// and for gfx950.
//
// The reference decides these forms in two stages (src/filter.rs:782-810,
// 1044-1125): what the key alone settles, and a prefilter that says whether a
// placement reaching into the checksum is still possible; only then does it
// pay for the sign of x and a SHA3-256.  Here both stages read one table:
//
//   * each pattern owns consecutive bits of a 32-bit state word (a "bank");
//     b[c] has bit (offset + i) set when symbol c is allowed at position i;
//   * per address symbol t_j:  D = ((D << 1) | inject_j) & b[t_j];  a match
//     ends at j when a pattern's last bit is set in D.
//
// Key pass (canonical y, sign clear): symbols 0..48 are exact, so a match
// ending there is a hit (src/filter.rs:120-134).  Symbols 49..55 continue as a
// prefilter: 49 accepts both values of the deferred sign bit (either_sign,
// src/filter.rs:158-164), 50 is exact, 51 has one known bit (y bit 248), 52
// and 53 come from the checksum, 54 is one of a/i/q/y and 55 is always d
// (src/key.rs:309-327).  A lane whose prefilter fires computes the sign and
// the checksum and runs the exact pass over all 56 symbols.
//
// Patterns longer than one bank (32 symbols) keep their first 32 positions;
// reaching the last of those marks the lane for the symbol-by-symbol check on
// the full address (`flong`), which such a match makes rare.
//
// Regular expressions ride on the same banks: each contributes its obligatory
// class sequences ("proxies", og_regex.hpp), whose last bits (`fmaybe`) never
// make a hit.  A proxy occurrence only sends the lane to the expression's
// deterministic automaton, which decides exactly: over symbols 0..48 first,
// then, if the expression is still open, over the whole address.
#pragma once
#include "fe.hpp"

namespace og {

constexpr int kBankBits = 32;
constexpr int kMaxBanks = 8;
constexpr int kBankWords = 52;

// One bank, 48 words (the device copy is these words back to back).
struct ScanBank {
  uint32_t b[32];   // per symbol value: the positions that accept it
  uint32_t istart;  // first bits of start-anchored patterns (injected at 0)
  uint32_t iany;    // first bits of the others (injected at every symbol)
  uint32_t fany;    // last bits that end a match anywhere
  uint32_t fend;    // last bits that count only at symbol 55
  uint32_t flong;   // last kept bits of truncated patterns
  uint32_t u51[2];  // symbol 51 by its known top bit (y bit 248)
  uint32_t uall;    // any symbol (52, 53)
  uint32_t u54;     // a, i, q or y
  uint32_t jstart;  // no pattern of the bank starts before this symbol
  uint32_t fmaybe;  // last bits of expression proxies
  // A bank whose patterns that may start inside the key are at most two
  // literals (the others end the address): its key pass compares them at
  // their placements inside symbols 0..48 instead of walking every symbol.
  // lit is their count; per literal, its first symbols' bits from bit 31 down
  // (lit_val) under lit_mask, and lit_info: bit j - 44 set where placement j
  // in 44..48 ends the literal by symbol 48, kLitStart (placement 0 only) and
  // kLitProxy (an expression's proxy: a match asks for its automata).
  uint32_t lit;
  uint32_t lit_val[2], lit_mask[2], lit_info[2];
  uint32_t pad[2];
};
constexpr uint32_t kLitStart = 1u << 8, kLitProxy = 1u << 9;
static_assert(sizeof(ScanBank) == kBankWords * 4, "bank layout");

// Symbol j of the address bit string held as big-endian words (eight for the
// key, nine with the checksum and version).
OG_HD uint32_t sym_at(const uint32_t *w, int j) {
  const int bit = 5 * j, i = bit >> 5, sh = bit & 31;
  if (sh <= 27) return (w[i] >> (27 - sh)) & 31u;
  return ((w[i] << (sh - 27)) | (w[i + 1] >> (59 - sh))) & 31u;
}

OG_HD void key_words(const Fe &c, uint32_t w[8]) {
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const uint32_t x = c.v[l];
    w[l] = (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
  }
}

// The key pass of one bank.  `hit`: a placement inside symbols 0..48
// matched.  `need`: some placement reaching past 48 is still possible (or a
// truncated pattern reached its last kept bit), so the full address decides.
// `maybe`: an expression proxy occurred, or may reach past 48.  Each flag is
// or-ed in, so the banks of a set accumulate.
// The 32 bits from bit POS on of a string held in N big-endian words (zeros
// past them).
template <int POS, int N>
OG_HD uint32_t key_bits32(const uint32_t *w) {
  constexpr int i = POS >> 5, sh = POS & 31;
  static_assert(i < N, "window past the string");
  if constexpr (sh == 0) {
    return w[i];
  } else {
    const uint32_t next = i + 1 < N ? w[i + 1 < N ? i + 1 : N - 1] : 0u;
    return (w[i] << sh) | (next >> (32 - sh));
  }
}

// The key pass of a literal bank, one literal: the least masked difference
// between the literal and the key at placements J..48 (m the least so far), so
// 0 where the literal sits.  Every placement up to 43 fits a literal of at
// most 6 symbols; placements 44..48 count where bit J - 44 of `tail` says the
// literal ends by symbol 48.  Three instructions a placement (a funnel shift,
// the masked difference, a minimum), with no table read and no branch.
template <int J>
OG_HD uint32_t literal_min_from(const uint32_t w[8], uint32_t val, uint32_t mask, uint32_t tail,
                                uint32_t m) {
  if constexpr (J > 48) {
    return m;
  } else {
    uint32_t x = (key_bits32<5 * J, 8>(w) ^ val) & mask;
    if constexpr (J >= 44) x |= ((tail >> (J - 44)) & 1u) ^ 1u;
#ifdef __HIP_DEVICE_COMPILE__
    // keep the scheduler from hoisting every placement's shift at once, which
    // would hold 49 values live in a loop already at its register limit
    if constexpr (J % 8 == 7) __builtin_amdgcn_sched_barrier(0);
#endif
    return literal_min_from<J + 1>(w, val, mask, tail, x < m ? x : m);
  }
}

OG_HD void scan_bank_key(const ScanBank &s, const uint32_t w[8], bool &hit, bool &need,
                         bool &maybe) {
  const uint32_t y248 = w[7] & 1u;  // stream bit 255: the top bit of symbol 51
  const uint32_t js = s.jstart, iany = s.iany, fm = s.fmaybe;
  uint32_t D = 0, acc = 0;
  if (s.lit) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if ((uint32_t)p < s.lit) {
        const uint32_t info = s.lit_info[p];
        const uint32_t m =
            (info & kLitStart) ? (key_bits32<0, 8>(w) ^ s.lit_val[p]) & s.lit_mask[p]
                               : literal_min_from<0>(w, s.lit_val[p], s.lit_mask[p], info, ~0u);
        if (info & kLitProxy) maybe |= m == 0;
        else hit |= m == 0;
      }
    }
    // D after symbol 48 from the partial matches a placement from 43 on
    // leaves (one of at most 7 symbols reaching past 48 starts at 43 or
    // later; a start-anchored one never does)
#pragma unroll
    for (int j = 43; j < 49; ++j) D = ((D << 1) | iany) & s.b[sym_at(w, j)];
  } else if (js < 49) {
    // A bank whose patterns all start past the key span skips the key pass.
    // Otherwise every symbol runs, with the injection masked before jstart
    // (no branch per symbol): D stays 0 until a pattern may start.
#pragma unroll
    for (int j = 0; j < 49; ++j) {
      const uint32_t inj = (j == 0 ? s.istart : 0u) | ((uint32_t)j >= js ? iany : 0u);
      D = ((D << 1) | inj) & s.b[sym_at(w, j)];
      acc |= D;
    }
  }
  hit |= (acc & s.fany) != 0;
  uint32_t e = acc & s.flong, m = acc;
  const uint32_t t49 = sym_at(w, 49);  // sign bit clear: value 2 of the symbol
  D = ((D << 1) | iany) & (s.b[t49] | s.b[t49 | 2u]);
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.b[sym_at(w, 50)];
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.u51[y248];
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.uall;
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.uall;
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.u54;
  e |= D & (s.fany | s.flong);
  m |= D;
  D = ((D << 1) | iany) & s.b[3];  // 'd'
  e |= D & (s.fany | s.flong | s.fend);
  m |= D;
  need |= e != 0;
  maybe |= (m & fm) != 0;
}

// The key pass over every bank of a set.  The first bank runs on its own:
// its table address is then a constant, so every symbol's read takes the
// table offset as an immediate instead of an address add (a set of one bank
// is the common case).
OG_HD void scan_key(const uint32_t w[8], const ScanBank *banks, uint32_t nbanks, bool &hit,
                    bool &need, bool &maybe) {
  hit = false;
  need = false;
  maybe = false;
  if (nbanks == 0) return;
  scan_bank_key(banks[0], w, hit, need, maybe);
  for (uint32_t k = 1; k < nbanks; ++k) scan_bank_key(banks[k], w, hit, need, maybe);
}

// The exact pass over the 35-byte address as nine big-endian words (the last
// holds checksum[0], checksum[1], version 3 and a zero byte).  Returns 1 for a
// match, 2 when only a truncated pattern's kept part matched (the caller
// checks those symbol by symbol), 0 otherwise.
OG_HD uint32_t scan_address(const uint32_t w[9], const ScanBank *banks, uint32_t nbanks) {
  uint32_t out = 0;
  for (uint32_t k = 0; k < nbanks; ++k) {
    const ScanBank &s = banks[k];
    const uint32_t js = s.jstart, iany = s.iany;
    uint32_t D = 0, acc = 0;
    if (js >= 48) {
      // every pattern of the bank starts in the last 8 symbols (a suffix of
      // at most 8): nothing before symbol 48 can change D
#pragma unroll
      for (int j = 48; j < 56; ++j) {
        D = ((D << 1) | ((uint32_t)j >= js ? iany : 0u)) & s.b[sym_at(w, j)];
        acc |= D;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 56; ++j) {
        const uint32_t inj = (j == 0 ? s.istart : 0u) | ((uint32_t)j >= js ? iany : 0u);
        D = ((D << 1) | inj) & s.b[sym_at(w, j)];
        acc |= D;
      }
    }
    if ((acc & s.fany) | (D & s.fend)) return 1;
    if (acc & s.flong) out = 2;
  }
  return out;
}

// The address words of a key: canonical y (sign clear), the sign of x and the
// two checksum bytes.
OG_HD void address_words(const uint32_t kw[8], uint32_t sign, uint32_t chk, uint32_t w[9]) {
#pragma unroll
  for (int l = 0; l < 8; ++l) w[l] = kw[l];
  w[7] |= sign << 7;  // the top bit of key byte 31
  w[8] = ((chk & 0xffu) << 24) | (((chk >> 8) & 0xffu) << 16) | (0x03u << 8);
}

// Regular expressions (og_regex.hpp): one deterministic automaton each, as
// [states][32] 16-bit entries: the next state, and in bit 15 whether the
// address may end there.  State 0 can no longer match, state 1 has matched.
// `recs` holds three words per expression: its entries' offset, its first
// state, and the word offset of its tail-liveness bits in `live` (bit 2s+b:
// from state s after symbol 50 a match can still end by symbol 55 when the
// top bit of symbol 51 is b).

// From state s after symbol 48 of an open expression: whether a match can
// still end in symbols 49..55.  Symbol 49 takes both values of the deferred
// sign bit, symbol 50 is exact, and the tail-liveness bit covers the rest.
// The table pointers are template parameters so that the device can pass
// them in the global address space (plain global loads, not flat ones).
template <class T16, class T32>
OG_HD bool regex_open(T16 t, T32 lv, uint32_t s, const uint32_t w[8]) {
  const uint32_t t49 = sym_at(w, 49), t50 = sym_at(w, 50), y248 = w[7] & 1u;
  for (uint32_t sign = 0; sign < 4; sign += 2) {
    const uint32_t e = t[s * 32u + (t49 | sign)];
    if (e == 1) return true;
    const uint32_t e2 = t[e * 32u + t50];
    const uint32_t bit = 2 * e2 + y248;
    if (e2 == 1 || ((lv[bit >> 5] >> (bit & 31)) & 1u)) return true;
  }
  return false;
}

// One walk of an automaton over symbols 0..48 from state s, reading the
// table at `t`.  TP is the pointer type: the device walks either the global
// table or, when the whole automaton fits, its LDS copy.  Table entries hold
// the next state alone (the end flags are a bit array of their own), so a
// step is the next row's offset or'ed with the symbol and one load.  The walk
// is unrolled, and each symbol is taken from the key words at a constant bit
// offset - a shift and a mask, or a funnel shift where it straddles two words
// - so it costs no register per symbol and no queue of words to shift along.
template <int J>
OG_HD uint32_t sym_c(const uint32_t *w) {
  constexpr int bit = 5 * J, i = bit >> 5, sh = bit & 31;
  if constexpr (sh <= 27) {
    return (w[i] >> (27 - sh)) & 31u;
  } else {
    return (uint32_t)((((uint64_t)w[i] << 32) | w[i + 1]) >> (59 - sh)) & 31u;
  }
}

// Left to itself the scheduler computes every symbol up front, which holds
// one register per symbol for the whole walk; a scheduling barrier after
// each step keeps each symbol next to the step that reads it.
OG_HD void step_barrier() {
#ifdef __HIP_DEVICE_COMPILE__
  __builtin_amdgcn_sched_barrier(0);
#endif
}

template <int J, int N, class TP>
OG_HD uint32_t walk_from(TP t, uint32_t s, const uint32_t *w) {
  if constexpr (J == N) {
    return s;
  } else {
    const uint32_t next = t[(s << 5) | sym_c<J>(w)];
    step_barrier();
    return walk_from<J + 1, N>(t, next, w);
  }
}

// Two walks step by step, so that each hides the other's table latency.
template <int J, int N, class TP>
OG_HD void walk2_from(TP t, uint32_t &a, uint32_t &b, const uint32_t *wa, const uint32_t *wb) {
  if constexpr (J < N) {
    a = t[(a << 5) | sym_c<J>(wa)];
    b = t[(b << 5) | sym_c<J>(wb)];
    step_barrier();
    walk2_from<J + 1, N>(t, a, b, wa, wb);
  }
}

// A table pointer every lane of the wave holds alike (each wave searches one
// filter set), moved to scalar registers, so that a step's load takes the
// table as its scalar base and the 32-bit row-and-symbol offset alone.
template <class TP>
OG_HD TP wave_uniform(TP p) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint64_t x = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return (TP)(((uint64_t)hi << 32) | lo);
#else
  return p;
#endif
}

// The key words as values of this walk: a walk inside a loop over automata
// would otherwise have its symbols taken once, before the loop, and held in
// registers through it.
template <int N>
OG_HD void fresh_words(const uint32_t *w, uint32_t out[N]) {
#pragma unroll
  for (int l = 0; l < N; ++l) {
    uint32_t x = w[l];
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(x));
#endif
    out[l] = x;
  }
}

template <class TP>
OG_HD uint32_t regex_walk49(TP t, uint32_t s, const uint32_t w[8]) {
  uint32_t f[8];
  fresh_words<8>(w, f);
  return walk_from<0, 49>(t, s, f);
}

template <class TP>
OG_HD void regex_walk49x2(TP t, uint32_t &sa, uint32_t &sb, const uint32_t wa[8],
                          const uint32_t wb[8]) {
  uint32_t fa[8], fb[8];
  fresh_words<8>(wa, fa);
  fresh_words<8>(wb, fb);
  walk2_from<0, 49>(t, sa, sb, fa, fb);
}

// The window prefilter of a dictionary of plain substrings (every word at
// least W symbols, unanchored, one symbol per position): two bitmaps of 2^18
// bits, one over the first 18 and one over the last 18 of the 5W bits of a
// W-symbol window.  Each word maps its windows at its symbols 0..S-1 (every
// word has at least W + S - 1 symbols), so a word at placement p shows in the
// window at the next multiple of S, which ends no later than the word does:
// probing placements 0, S, 2S, ... finds every word, and a string whose
// probed windows all miss one of the maps holds none.  W is 5, 6 or 7 and S
// 1 to 4, packed as 8 W + S (DevFilterArgs::hash_ws).
constexpr uint32_t kHashSliceBits = 18;
constexpr uint32_t kHashSliceWords = 1u << (kHashSliceBits - 5);  // per map

// 32 bits whose low 18 are the string's bits P..P+17 (a map slot: its word
// is bits 5..17, its bit bits 0..4, which a shift reads by itself).
template <int P, int N>
OG_HD uint32_t slice_bits(const uint32_t *w) {
  constexpr int kUp = 32 - (int)kHashSliceBits;
  if constexpr (P >= kUp) return key_bits32<P - kUp, N>(w);
  else return w[0] >> (kUp - P);
}

// The bit of map word `word` a slice names (the shift takes its low 5 bits).
OG_HD uint32_t slot_bit(uint32_t word, uint32_t slice) { return word >> (slice & 31u); }

// Whether a window at placements J, J + S, ... <= LAST hits both maps.
template <int W, int S, int LAST, int N, int J, class TL>
OG_HD uint32_t window_hits_from(const uint32_t *w, TL maps) {
  if constexpr (J > LAST) {
    return 0u;
  } else {
    constexpr uint32_t kWords = kHashSliceWords - 1;
    const uint32_t a = slice_bits<5 * J, N>(w);
    const uint32_t b = slice_bits<5 * (J + W) - kHashSliceBits, N>(w);
    const uint32_t hit = slot_bit(maps[(a >> 5) & kWords], a) &
                         slot_bit(maps[kHashSliceWords + ((b >> 5) & kWords)], b) & 1u;
    return hit | window_hits_from<W, S, LAST, N, J + S>(w, maps);
  }
}

// The windows wholly inside the first END symbols of a string of N words
// (END 49 with the 8 key words: the symbols the key settles; END 56 with the
// 9 address words: the whole address).
template <int END, int N, class TL>
OG_HD bool window_hit(const uint32_t *w, TL maps, uint32_t ws) {
  switch (ws) {
#define OG_WINDOW_CASE(W, S) \
  case 8 * W + S: return window_hits_from<W, S, END - W, N, 0>(w, maps) != 0;
    OG_WINDOW_CASE(5, 1) OG_WINDOW_CASE(5, 2) OG_WINDOW_CASE(5, 3) OG_WINDOW_CASE(5, 4)
    OG_WINDOW_CASE(6, 1) OG_WINDOW_CASE(6, 2) OG_WINDOW_CASE(6, 3) OG_WINDOW_CASE(6, 4)
    OG_WINDOW_CASE(7, 1) OG_WINDOW_CASE(7, 2) OG_WINDOW_CASE(7, 3) OG_WINDOW_CASE(7, 4)
#undef OG_WINDOW_CASE
    default: return true;  // no such form: every string goes on to the exact pass
  }
}

// The states after symbols j0..48 for the two keys of a pair, from the
// states given (the short walk of the prefiltered dictionaries, with a
// run-time first symbol), their steps interleaved so that each hides the
// other's table latency.
template <class TP>
OG_HD void walk_tail2(TP t, uint32_t &sa, uint32_t &sb, const uint32_t wa[8],
                      const uint32_t wb[8], uint32_t j0) {
  uint32_t fa[8], fb[8];
  fresh_words<8>(wa, fa);
  fresh_words<8>(wb, fb);
  switch (j0) {  // the common word lengths 8..12 take the unrolled steps
    case 42: walk2_from<42, 49>(t, sa, sb, fa, fb); return;
    case 41: walk2_from<41, 49>(t, sa, sb, fa, fb); return;
    case 40: walk2_from<40, 49>(t, sa, sb, fa, fb); return;
    case 39: walk2_from<39, 49>(t, sa, sb, fa, fb); return;
    case 38: walk2_from<38, 49>(t, sa, sb, fa, fb); return;
    default:
      for (uint32_t j = j0; j < 49; ++j) {
        sa = t[(sa << 5) | sym_at(fa, (int)j)];
        sb = t[(sb << 5) | sym_at(fb, (int)j)];
      }
  }
}

// Symbols 0..48 (what the key settles) through every expression: 2 when one
// has matched, 1 when one may still match on the address, 0 otherwise.
// `in_lds`: the first expression's whole table is also at `lds`.
template <class T16, class T32, class TL>
OG_HD uint32_t regex_key(const uint32_t w[8], const uint32_t *recs, T16 dfa, T32 live,
                         uint32_t n, TL lds, uint32_t in_lds) {
  uint32_t out = 0;
  for (uint32_t r = 0; r < n; ++r) {
    const T16 t = dfa + recs[3 * r];
    uint32_t s = recs[3 * r + 1];
    s = (r == 0 && in_lds) ? regex_walk49(lds, s, w) : regex_walk49(wave_uniform(t), s, w);
    if (s == 1) return 2;
    if (s != 0 && !out && regex_open(t, live + recs[3 * r + 2], s, w)) out = 1;
  }
  return out;
}

// The same for two keys at once, their walks interleaved so that each hides
// the other's table latency: the first key's verdict in bits 0-1, the
// second's in bits 2-3.
template <class T16, class T32, class TL>
OG_HD uint32_t regex_key2(const uint32_t wa[8], const uint32_t wb[8], const uint32_t *recs,
                          T16 dfa, T32 live, uint32_t n, TL lds, uint32_t in_lds) {
  uint32_t ra = 0, rb = 0;
  for (uint32_t r = 0; r < n && (ra != 2 || rb != 2); ++r) {
    const T16 t = dfa + recs[3 * r];
    const T32 lv = live + recs[3 * r + 2];
    uint32_t sa = recs[3 * r + 1], sb = sa;
    if (r == 0 && in_lds) regex_walk49x2(lds, sa, sb, wa, wb);
    else regex_walk49x2(wave_uniform(t), sa, sb, wa, wb);
    if (sa == 1) ra = 2;
    else if (sa != 0 && !ra && regex_open(t, lv, sa, wa)) ra = 1;
    if (sb == 1) rb = 2;
    else if (sb != 0 && !rb && regex_open(t, lv, sb, wb)) rb = 1;
  }
  return ra | (rb << 2);
}

// The key stage of a dictionary of plain substrings behind the window
// prefilter (`maps`, its form `ws`, the longest word `lmax`): 2 when a word has
// matched, 1 when the address stage must decide, 0 when no word can occur.
// A word that ends at symbol 49 or later starts at 50 - lmax or later, and the
// automaton of unanchored words forgets everything older than its longest
// word, so a walk from the start state over symbols 50-lmax..48 reaches the
// state the whole walk reaches unless a word ended earlier.  Such a word,
// wholly inside symbols 0..48, shows in a window that hits both maps, and a
// key with such a window goes to the address stage (1), whose exact pass
// over the whole address finds it: parked candidates are compacted per wave,
// so the rare window hits cost one address stage each instead of a whole
// walk for every lane of their waves.  The verdict differs from regex_key2's
// only by 1 in place of 0 or 2 where a window hits, which the address stage
// settles.
template <class T16, class T32, class TL>
OG_HD uint32_t regex_key2_prefiltered(const uint32_t wa[8], const uint32_t wb[8],
                                      const uint32_t *recs, T16 dfa, T32 live, uint32_t n,
                                      TL maps, uint32_t lmax, uint32_t ws) {
  const uint32_t j0 = lmax < 49 ? 50 - lmax : 1;
  const bool pa = window_hit<49, 8>(wa, maps, ws), pb = window_hit<49, 8>(wb, maps, ws);
  uint32_t ra = pa ? 1u : 0u, rb = pb ? 1u : 0u;
  for (uint32_t r = 0; r < n && (ra != 2 || rb != 2); ++r) {
    const T16 t = wave_uniform(dfa + recs[3 * r]);
    const T32 lv = live + recs[3 * r + 2];
    uint32_t sa = recs[3 * r + 1], sb = sa;
    walk_tail2(t, sa, sb, wa, wb, j0);
    if (sa == 1) ra = 2;
    else if (sa != 0 && !ra && regex_open(t, lv, sa, wa)) ra = 1;
    if (sb == 1) rb = 2;
    else if (sb != 0 && !rb && regex_open(t, lv, sb, wb)) rb = 1;
  }
  return ra | (rb << 2);
}

// All 56 symbols of the address words: whether an expression matches.  The
// last symbol's entry is read from the end flags: whether the address may
// end there.
template <class T16, class T32>
OG_HD bool regex_address(const uint32_t w[9], const uint32_t *recs, T16 dfa, T32 ends,
                         uint32_t n) {
  for (uint32_t r = 0; r < n; ++r) {
    uint32_t f[9];
    fresh_words<9>(w, f);
    const uint32_t s = walk_from<0, 55>(wave_uniform(dfa + recs[3 * r]), recs[3 * r + 1], f);
    if ((ends[(recs[3 * r] >> 5) + s] >> sym_c<55>(w)) & 1u) return true;
  }
  return false;
}

// The exact table behind a bitmap: per distinct prefix length a bucket of
// kBucketWords words - offset and count of its sorted keys (masked big-endian
// 64-bit heads), the mask (low, high word), its directory's offset and the
// shift that takes a key's top D bits as the directory index.  The directory
// has 2^D + 1 entries: entry t is the first key whose top D bits are t or more.
constexpr int kBucketWords = 8;
constexpr uint32_t kDirMaxBits = 20;

// The exact table: whether the key's first 64 bits (big-endian, head_hi
// first), masked to some bucket's prefix length, equal one of that bucket's
// sorted keys.  Four buckets at a time side by side, so their loads are in
// flight together on the device: the directory entry of the key's top bits gives the run of
// keys that share them (about one), and `steps` halvings of the run find the
// first key not below the probe.  Out of line: the rare lanes that get here
// (a bitmap pass or an inexact probe word) keep its registers out of the
// search loop's.  The host runs the same code for the tests.
static __host__ __device__ __noinline__ bool exact_table(uint32_t head_hi, uint32_t head_lo,
                                         const uint32_t *buckets, const uint64_t *keys,
                                         const uint32_t *dirs, uint32_t nbuckets, uint32_t steps) {
  const uint64_t h = ((uint64_t)head_hi << 32) | head_lo;
  bool found = false;
  for (uint32_t b0 = 0; b0 < nbuckets; b0 += 4) {
    uint64_t probe[4];
    uint32_t lo[4], hi[4], end[4], off[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t b = b0 + j < nbuckets ? b0 + j : nbuckets - 1;  // wave-uniform
      const uint32_t *r = buckets + (size_t)kBucketWords * b;
      probe[j] = h & ((uint64_t)r[2] | ((uint64_t)r[3] << 32));
      const uint32_t t = (uint32_t)(probe[j] >> r[5]);
      off[j] = r[0];
      lo[j] = dirs[r[4] + t];
      hi[j] = end[j] = b0 + j < nbuckets ? dirs[r[4] + t + 1] : lo[j];
    }
    for (uint32_t s = 0; s < steps; ++s) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t mid = (lo[j] + hi[j]) >> 1;
        if (lo[j] < hi[j]) {
          if (keys[off[j] + mid] < probe[j]) lo[j] = mid + 1;
          else hi[j] = mid;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (lo[j] < end[j]) found |= keys[off[j] + lo[j]] == probe[j];
  }
  return found;
}

}  // namespace og
