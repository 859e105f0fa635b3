// Device side of the search, shared by the single-search engine and the
// batched (slot) engine: SoA access, the filter tests, the wave-level hit
// append and the window loop.  See og_kernels.hpp for the method.
#pragma once
#include "og_kernels.hpp"

namespace og {

typedef const __attribute__((address_space(4))) uint32_t *cptr;

// Occupancy of the search kernels, in waves per SIMD.  With each product
// column summed as one dependent chain (fe10_mul) the key forms' loop fits 168
// VGPRs and runs three waves per SIMD (measured 2% over two with the same
// products); the pattern forms carry the automata's registers as well and stay
// at two (256 VGPRs), where three spill inside the loop and measured 5-25%
// slower.
#ifndef OG_SEARCH_WAVES_KEY
#define OG_SEARCH_WAVES_KEY 3
#endif
#ifndef OG_SEARCH_WAVES_PATTERN
#define OG_SEARCH_WAVES_PATTERN 2
#endif
template <uint32_t MODE>
constexpr unsigned kSearchWaves =
    (MODE & kModePattern) ? OG_SEARCH_WAVES_PATTERN : OG_SEARCH_WAVES_KEY;
#define OG_SEARCH_ATTR(MODE) \
  __attribute__((amdgpu_waves_per_eu(kSearchWaves<MODE>, kSearchWaves<MODE>)))

__device__ __forceinline__ Fe10 ld10_const(cptr p) {
  Fe10 r;
#pragma unroll
  for (int l = 0; l < 10; ++l) r.v[l] = p[l];
  return r;
}

// Element `plane` of a structure-of-arrays buffer of 10-limb values: limb l of
// chain g sits at (plane*10 + l)*T + g, so a wave touches 64 consecutive words.
// Each access is a raw buffer op: the plane base in a scalar descriptor (built
// by SALU, wave-uniform), the limb offset l*T*4 as the scalar offset and g*4 as
// the one lane offset, so no load or store spends a VALU slot on 64-bit
// address arithmetic.  T is at most 2^22 (both engines refuse more), so every
// in-plane offset stays below 40 * 2^22 bytes, inside the descriptor's range.
constexpr int kBufWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format
constexpr int kBufNT = 2;              // the nontemporal cache policy bit

__device__ __forceinline__ __amdgpu_buffer_rsrc_t soa_plane(const uint32_t *base, size_t plane,
                                                            uint32_t T) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(base + plane * 10 * T), 0,
                                           0x7fffffff, kBufWord3);
}

__device__ __forceinline__ Fe10 ld10(const uint32_t *base, size_t plane, uint32_t T, uint32_t g) {
  const __amdgpu_buffer_rsrc_t r = soa_plane(base, plane, T);
  Fe10 v;
#pragma unroll
  for (int l = 0; l < 10; ++l)
    v.v[l] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(g * 4u), (int)(l * T * 4u), kBufNT);
  return v;
}

__device__ __forceinline__ void st10(uint32_t *base, size_t plane, uint32_t T, uint32_t g,
                                     const Fe10 &v) {
  const __amdgpu_buffer_rsrc_t r = soa_plane(base, plane, T);
#pragma unroll
  for (int l = 0; l < 10; ++l)
    __builtin_amdgcn_raw_buffer_store_b32(v.v[l], r, (int)(g * 4u), (int)(l * T * 4u), kBufNT);
}

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

__device__ __forceinline__ bool bitmap_bit(const DevFilterArgs &f, uint32_t head) {
  const uint32_t slot = (uint32_t)((uint64_t)head >> (32 - f.bits));
  return (f.bitmap[slot >> 5] >> (slot & 31)) & 1u;
}

// The cheap test every candidate gets: on the low bits of the key (its first
// three bytes), taken from three columns of its product (fe10_mul_low26).
// Never false for a candidate the exact test accepts: where that word is not
// settled the probe tests each value the canonical word can take.
// A MODE with both key bits set leaves the key form to the filter at run time
// (f.mode, wave-uniform): one instantiation serves every key form, which keeps
// the slot engine's kernels to one search loop each.
template <uint32_t MODE>
__device__ __forceinline__ bool key_is_prefix(const DevFilterArgs &f) {
  constexpr bool kBoth = (MODE & kModePrefix) && (MODE & kModeBitmap);
  return kBoth ? (f.mode & kModePrefix) != 0 : (MODE & kModePrefix) != 0;
}

template <uint32_t MODE>
__device__ __forceinline__ bool key_is_bitmap(const DevFilterArgs &f) {
  constexpr bool kBoth = (MODE & kModePrefix) && (MODE & kModeBitmap);
  return kBoth ? (f.mode & kModeBitmap) != 0 : (MODE & kModeBitmap) != 0;
}

// The bitmap probe through the workgroup's LDS copy (`coarse`, null when the
// kernel keeps none): the coarse map first, then the full map in L2 for the
// lanes whose coarse bit is set, unless the copy is the full map itself.
__device__ __forceinline__ bool coarse_bit(const uint32_t *coarse, uint32_t cbits, uint32_t head) {
  const uint32_t slot = head >> (32 - cbits);
  return (coarse[slot >> 5] >> (slot & 31)) & 1u;
}

// The probe word of a*b for the key forms: bits 0..25 of the canonical product
// from three columns (the first three key bytes, all a prefix compare or a
// map of up to 24 bits reads), with column 1 added for a wider map.
template <uint32_t MODE>
__device__ __forceinline__ uint32_t key_probe_word(const Fe10 &a, const Fe10 &b,
                                                   const DevFilterArgs &f, bool &exact) {
  uint32_t lo = fe10_mul_low26(a, b, exact);
  if (key_is_bitmap<MODE>(f) && f.bits > 24) lo += fe10_low_col1(a, b);  // wave-uniform
  return lo;
}

// Whether the map holds the slot of probe word x: the workgroup's LDS fold
// first, then the map in L2 for the lanes whose fold bit is set (the others
// read word 0, one line for all of them).
template <uint32_t MODE>
__device__ __forceinline__ uint32_t map_bit(uint32_t x, const DevFilterArgs &f,
                                            const uint32_t *coarse) {
  const uint32_t h = bswap32(x);
  if (!coarse) return (uint32_t)bitmap_bit(f, h);
  const uint32_t c = (uint32_t)coarse_bit(coarse, f.coarse_bits, h);
  if (f.coarse_bits == f.bits) return c;
  const uint32_t slot = (uint32_t)((uint64_t)h >> (32 - f.bits));
  const uint32_t word = c ? slot >> 5 : 0u;
  return c & (f.bitmap[word] >> (slot & 31)) & 1u;
}

// Where the probe word is inexact the canonical word is lo, lo + 19 or
// lo + 38, so those lanes test all three, in the waves that hold such a lane
// (a wave-uniform branch, taken for about one candidate slot in 16): only a
// lane where one of them passes goes on to the whole product.
template <uint32_t MODE>
__device__ __forceinline__ bool maybe_key(const Fe10 &a, const Fe10 &b, const DevFilterArgs &f,
                                          const uint32_t *coarse = nullptr) {
  bool exact;
  const uint32_t lo = key_probe_word<MODE>(a, b, f, exact);
  if (key_is_prefix<MODE>(f)) {
    // a prefix compares its first three bytes here; settle_key compares the rest
    const uint32_t m = f.mask[0] & kLow26, want = f.want[0];
    bool pass = ((lo ^ want) & m) == 0;
    if (__ballot(!exact))
      pass |= !exact && ((((lo + 19u) ^ want) & m) == 0 || (((lo + 38u) ^ want) & m) == 0);
    return pass;
  }
  if (key_is_bitmap<MODE>(f)) {
    uint32_t pass = map_bit<MODE>(lo, f, coarse);
    if (__ballot(!exact) && !exact)
      pass |= map_bit<MODE>(lo + 19u, f, coarse) | map_bit<MODE>(lo + 38u, f, coarse);
    return pass != 0;
  }
  return false;
}

// The exact key-side test on the canonical key (sign clear).
template <uint32_t MODE>
__device__ __forceinline__ bool exact_key(const Fe &c, const DevFilterArgs &f) {
  if (key_is_prefix<MODE>(f)) {
    uint32_t diff = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) diff |= (c.v[l] ^ f.want[l]) & f.mask[l];
    return diff == 0;
  }
  if (key_is_bitmap<MODE>(f)) {
    const uint32_t head = bswap32(c.v[0]);
    // the table holds every key the map does, so with one it alone decides
    if (f.nbuckets == 0) return bitmap_bit(f, head);
    return exact_table(head, bswap32(c.v[1]), f.buckets, f.keys, f.dirs, f.nbuckets, f.dir_steps);
  }
  return false;
}

// The general forms symbol by symbol on the full address: the exact check for
// patterns longer than one automaton bank.  Out of line, with its own copy of
// the address, so that its run-time symbol index does not push the caller's
// words into scratch memory.
__device__ __noinline__ bool pattern_slow(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3,
                                          uint32_t w4, uint32_t w5, uint32_t w6, uint32_t w7,
                                          uint32_t w8, const uint32_t *pats, uint32_t npat) {
  const uint32_t w[9] = {w0, w1, w2, w3, w4, w5, w6, w7, w8};
  for (uint32_t p = 0; p < npat; ++p) {
    const uint32_t *rec = pats + p * kPatWords;
    const uint32_t len = rec[0], first = rec[1], last = rec[2];
    for (uint32_t s = first; s <= last; ++s) {
      bool ok = true;
      for (uint32_t i = 0; i < len && ok; ++i) ok = (rec[4 + i] >> sym_at(w, (int)(s + i))) & 1u;
      if (ok) return true;
    }
  }
  return false;
}

// The expressions' automata, out of line like pattern_slow: over the key's
// symbols 0..48 (2 matched, 1 open, 0 closed) and over the whole address.
// Their tables are read as global memory (hipMalloc'd by the host).
typedef const __attribute__((address_space(1))) uint16_t *gptr16;
typedef const __attribute__((address_space(1))) uint32_t *gptr32;
typedef const __attribute__((address_space(3))) uint16_t *lptr16;
typedef const __attribute__((address_space(3))) uint32_t *lptr32;
__device__ __noinline__ uint32_t regex_key_slow(uint32_t w0, uint32_t w1, uint32_t w2,
                                                uint32_t w3, uint32_t w4, uint32_t w5,
                                                uint32_t w6, uint32_t w7, const uint32_t *recs,
                                                const uint16_t *dfa, const uint32_t *live,
                                                uint32_t n, lptr16 lds, uint32_t K) {
  const uint32_t w[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
  return regex_key(w, recs, (gptr16)dfa, (gptr32)live, n, lds, K);
}

__device__ __noinline__ uint32_t regex_key2_slow(uint32_t a0, uint32_t a1, uint32_t a2,
                                                 uint32_t a3, uint32_t a4, uint32_t a5,
                                                 uint32_t a6, uint32_t a7, uint32_t b0,
                                                 uint32_t b1, uint32_t b2, uint32_t b3,
                                                 uint32_t b4, uint32_t b5, uint32_t b6,
                                                 uint32_t b7, const uint32_t *recs,
                                                 const uint16_t *dfa, const uint32_t *live,
                                                 uint32_t n, lptr16 lds, uint32_t K) {
  const uint32_t wa[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
  const uint32_t wb[8] = {b0, b1, b2, b3, b4, b5, b6, b7};
  return regex_key2(wa, wb, recs, (gptr16)dfa, (gptr32)live, n, lds, K);
}

// The dictionary of plain substrings behind the window prefilter, whose maps
// are in this workgroup's LDS.
__device__ __noinline__ uint32_t regex_key2_pref_slow(uint32_t a0, uint32_t a1, uint32_t a2,
                                                      uint32_t a3, uint32_t a4, uint32_t a5,
                                                      uint32_t a6, uint32_t a7, uint32_t b0,
                                                      uint32_t b1, uint32_t b2, uint32_t b3,
                                                      uint32_t b4, uint32_t b5, uint32_t b6,
                                                      uint32_t b7, const uint32_t *recs,
                                                      const uint16_t *dfa, const uint32_t *live,
                                                      uint32_t n, lptr32 maps, uint32_t lmax,
                                                      uint32_t ws) {
  const uint32_t wa[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
  const uint32_t wb[8] = {b0, b1, b2, b3, b4, b5, b6, b7};
  return regex_key2_prefiltered(wa, wb, recs, (gptr16)dfa, (gptr32)live, n, maps, lmax, ws);
}

__device__ __noinline__ bool regex_address_slow(uint32_t w0, uint32_t w1, uint32_t w2,
                                                uint32_t w3, uint32_t w4, uint32_t w5,
                                                uint32_t w6, uint32_t w7, uint32_t w8,
                                                const uint32_t *recs, const uint16_t *dfa,
                                                const uint32_t *ends, uint32_t n) {
  const uint32_t w[9] = {w0, w1, w2, w3, w4, w5, w6, w7, w8};
  return regex_address(w, recs, (gptr16)dfa, (gptr32)ends, n);
}

// Appends a wave's hits with one atomic per wave.  Records past `cap` are
// counted but not written: the host reports the overflow.
__device__ __forceinline__ void append(bool hit, uint32_t slot, uint64_t k, const Fe &c,
                                       uint32_t *counter, HitRec *hits, uint32_t cap) {
  const uint64_t mask = __ballot(hit);
  if (mask == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  if (hit) {
    const uint32_t idx = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
    if (idx < cap) {
      hits[idx].k = k;
      hits[idx].slot = slot;
#pragma unroll
      for (int l = 0; l < 8; ++l) hits[idx].y[l] = c.v[l];
    }
  }
}

struct HitSink {
  uint32_t *counter;
  HitRec *hits;
  uint32_t cap;
  uint32_t slot;
  uint32_t *found;   // the search's count of every hit this launch, or null
  uint32_t budget;   // with `found`: records the search may still write
};

// A hit is always counted; with a budget only the first `budget` of the
// search's hits this launch (in the order the waves count them) are written to
// the ring, so a search that wants n hits never floods it.
__device__ __forceinline__ void report(bool hit, uint64_t k, const Fe &c, const HitSink &sink) {
  if (!sink.found) {
    append(hit, sink.slot, k, c, sink.counter, sink.hits, sink.cap);
    return;
  }
  const uint64_t mask = __ballot(hit);
  if (mask == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(sink.found, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  const uint32_t rank = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
  append(hit && rank < sink.budget, sink.slot, k, c, sink.counter, sink.hits, sink.cap);
}

// Key forms only, for the candidate whose y is a*b and whose probe said
// `maybe` (maybe_key): the whole product and the exact test on the lanes the
// wave ballot names.
template <uint32_t MODE>
__device__ __forceinline__ void settle_key(bool maybe, const Fe10 &a, const Fe10 &b, uint64_t k,
                                           const DevFilterArgs &f, const HitSink &sink) {
  if (__ballot(maybe) == 0) return;  // the common case: no lane needs more
  bool hit = false;
  Fe c;
  if (maybe) {
    c = fe10_canon_product(fe10_mul(a, b));
    hit = exact_key<MODE>(c, f);
    if (f.key_proxy && hit) {
      // an expression's prefix: its automata settle the key's 49 symbols
      uint32_t w[8];
      key_words(c, w);
      hit = regex_key_slow(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], f.regex, f.dfa,
                           f.regex_live, f.nregex, nullptr, 0u) == 2;
    }
  }
  report(hit, k, c, sink);
}

// The low-bits probe of one candidate (three columns of its product), then
// settle_key.
template <uint32_t MODE>
__device__ __forceinline__ void test_key_only(const Fe10 &a, const Fe10 &b, uint64_t k,
                                              const DevFilterArgs &f, const HitSink &sink,
                                              const uint32_t *coarse = nullptr) {
  settle_key<MODE>(maybe_key<MODE>(a, b, f, coarse), a, b, k, f, sink);
}

// Both candidates of a pair, their probes side by side under one test of
// the set's wave-uniform form, so that a bitmap's reads (LDS, then L2) for
// the two are in flight together; then each is settled.
template <uint32_t MODE>
__device__ __forceinline__ void test_key_pair(const Fe10 &am, const Fe10 &ap, const Fe10 &b,
                                              uint64_t km, uint64_t kp, const DevFilterArgs &f,
                                              const HitSink &sink, const uint32_t *coarse) {
  bool mm, mp;
  if (key_is_bitmap<MODE>(f) && coarse && f.coarse_bits != f.bits) {
    bool em, ep;
    const uint32_t lm = key_probe_word<MODE>(am, b, f, em);
    const uint32_t lp = key_probe_word<MODE>(ap, b, f, ep);
    const uint32_t hm = bswap32(lm), hp = bswap32(lp);
    // every lane reads; a lane whose coarse bit is clear reads word 0 of the
    // map in L2, one line for all of them
    const uint32_t cm = (uint32_t)coarse_bit(coarse, f.coarse_bits, hm);
    const uint32_t cp = (uint32_t)coarse_bit(coarse, f.coarse_bits, hp);
    const uint32_t sm = (uint32_t)((uint64_t)hm >> (32 - f.bits));
    const uint32_t sp = (uint32_t)((uint64_t)hp >> (32 - f.bits));
    const uint32_t wm = f.bitmap[cm ? sm >> 5 : 0u], wp = f.bitmap[cp ? sp >> 5 : 0u];
    uint32_t pm = cm & (wm >> (sm & 31)) & 1u, pp = cp & (wp >> (sp & 31)) & 1u;
    // inexact words: the canonical word is also possibly lo + 19 or lo + 38
    if (__ballot(!em || !ep)) {
      if (!em) pm |= map_bit<MODE>(lm + 19u, f, coarse) | map_bit<MODE>(lm + 38u, f, coarse);
      if (!ep) pp |= map_bit<MODE>(lp + 19u, f, coarse) | map_bit<MODE>(lp + 38u, f, coarse);
    }
    mm = pm != 0;
    mp = pp != 0;
  } else {
    mm = maybe_key<MODE>(am, b, f, coarse);
    mp = maybe_key<MODE>(ap, b, f, coarse);
  }
  settle_key<MODE>(mm, am, b, km, f, sink);
  settle_key<MODE>(mp, ap, b, kp, f, sink);
}

// The key-derived part of a candidate under the general forms (and the key
// forms when the set has both): `hit` is final, `need` asks for the address.
// Lanes where an expression's proxy occurred (all lanes, for an expression
// without one) run the automata over symbols 0..48: matched is a hit, still
// open asks for the address.
//
// A set whose key stage would send nearly every candidate on (park_all)
// skips it: the key forms decide their hits, and every other candidate goes
// to the address stage, whose exact pass covers all 56 symbols.
template <uint32_t MODE>
__device__ __forceinline__ void key_stage(const Fe &c, const uint32_t w[8], const DevFilterArgs &f,
                                          const ScanBank *banks, const uint16_t *dfa_lds,
                                          bool &hit, bool &need) {
  if (f.park_all) {
    hit = (MODE & (kModePrefix | kModeBitmap)) ? exact_key<MODE>(c, f) : false;
    need = !hit;
    return;
  }
  bool maybe;
  scan_key(w, banks, f.nbanks, hit, need, maybe);
  if (MODE & (kModePrefix | kModeBitmap)) hit |= exact_key<MODE>(c, f);
  if (f.nregex) {
    maybe = (maybe || f.regex_every != 0) && !hit;
    if (__ballot(maybe) && maybe) {
      const uint32_t r =
          regex_key_slow(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], f.regex, f.dfa,
                         f.regex_live, f.nregex, (lptr16)dfa_lds, f.lds_states);
      hit |= r == 2;
      need |= r == 1;
    }
  }
  need &= !hit;
}

// key_stage for the two candidates of a pair, with one interleaved walk of
// the automata for both.
template <uint32_t MODE>
__device__ __forceinline__ void key_stage_pair(const Fe &ca, const uint32_t wa[8], const Fe &cb,
                                               const uint32_t wb[8], const DevFilterArgs &f,
                                               const ScanBank *banks, const uint16_t *dfa_lds,
                                               bool &ha, bool &na, bool &hb, bool &nb) {
  if (f.park_all) {
    ha = hb = false;
    if (MODE & (kModePrefix | kModeBitmap)) {
      ha = exact_key<MODE>(ca, f);
      hb = exact_key<MODE>(cb, f);
    }
    na = !ha;
    nb = !hb;
    return;
  }
  bool ma, mb;
  scan_key(wa, banks, f.nbanks, ha, na, ma);
  scan_key(wb, banks, f.nbanks, hb, nb, mb);
  if (MODE & (kModePrefix | kModeBitmap)) {
    ha |= exact_key<MODE>(ca, f);
    hb |= exact_key<MODE>(cb, f);
  }
  if (f.nregex) {
    ma = (ma || f.regex_every != 0) && !ha;
    mb = (mb || f.regex_every != 0) && !hb;
    if (__ballot(ma || mb) && (ma || mb)) {
      const uint32_t r =
          f.hash_key
              ? regex_key2_pref_slow(wa[0], wa[1], wa[2], wa[3], wa[4], wa[5], wa[6], wa[7],
                                     wb[0], wb[1], wb[2], wb[3], wb[4], wb[5], wb[6], wb[7],
                                     f.regex, f.dfa, f.regex_live, f.nregex,
                                     (lptr32)dfa_lds, f.hash_lmax, f.hash_ws)
              : regex_key2_slow(wa[0], wa[1], wa[2], wa[3], wa[4], wa[5], wa[6], wa[7], wb[0],
                                wb[1], wb[2], wb[3], wb[4], wb[5], wb[6], wb[7], f.regex, f.dfa,
                                f.regex_live, f.nregex, (lptr16)dfa_lds, f.lds_states);
      const uint32_t ra = ma ? (r & 3u) : 0u, rb = mb ? (r >> 2) : 0u;
      ha |= ra == 2;
      na |= ra == 1;
      hb |= rb == 2;
      nb |= rb == 1;
    }
  }
  na &= !ha;
  nb &= !hb;
}

// The address stage is deferred.  A lane whose key pass leaves a checksum
// placement possible computes the sign of x (bit 0 of its canonical form) and
// parks the candidate - packed key with the sign in bit 255, and its index -
// in a scratch plane the backward pass has already consumed, counting down
// from plane 3H.  After the backward pass the parked
// candidates are drained (SHA3-256, the exact automaton pass) with little
// else live, so the Keccak state never competes with the chain's.
//
// Parking is compacted per wave.  The wave keeps one count Q; the lanes that
// park at a step take entries Q, Q+1, ... in lane order, and entry q lands in
// plane 3H - floor(q/64), in the column of the wave's lane q mod 64.  So one
// step's stores are contiguous (lanes' own counts would diverge within a few
// steps and scatter every store over 64 planes), and the drain gives every
// lane the same share: lane i takes entries i, i+64, ..., in planes 3H, 3H-1,
// ... of its own column.  After the centre and pairs H-1..m the wave has at
// most 64(2(H-m)+1) entries, in at most 2(H-m)+1 planes, the lowest of them
// 3H - 2(H-m) = H + 2m, above the planes 0 .. 2m-1 of the pairs the backward
// pass has yet to read, and never below plane H.  `mask` is the wave's ballot of `want`.  `sign` is the sign of
// x, or 0 when the address kernel is to derive it (DevFilterArgs::defer_sign).
__device__ __forceinline__ void park(uint32_t *scratch, uint32_t top, uint32_t T, uint32_t g,
                                     uint32_t &Q, uint64_t mask, bool want, const Fe &c,
                                     uint32_t sign, uint64_t k) {
  if (want) {
    const uint32_t lane = __lane_id();
    const uint32_t q = Q + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
    Fe10 e;
#pragma unroll
    for (int l = 0; l < 8; ++l) e.v[l] = c.v[l];
    e.v[7] |= sign << 31;
    e.v[8] = (uint32_t)k;
    e.v[9] = (uint32_t)(k >> 32);
    st10(scratch, top - (q >> 6), T, g - lane + (q & 63u), e);
  }
  Q += (uint32_t)__popcll(mask);
}

// A lane's share of its wave's Q parked entries.
__device__ __forceinline__ uint32_t parked_share(uint32_t Q) {
  const uint32_t lane = __lane_id();
  return Q > lane ? (Q - lane + 63u) >> 6 : 0u;
}

// The address stage of one parked entry: packed key (sign in bit 255) and index.
// Behind the address stage's window prefilter (hash_addr, its maps at `maps`
// in LDS) the automata walk only where a window of the 56 symbols hits: a
// wave walks when one of its lanes needs it.
__device__ __forceinline__ void drain_one(const Fe10 &e, const DevFilterArgs &f,
                                          const ScanBank *banks, lptr32 maps,
                                          const HitSink &sink) {
  uint32_t pk[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) pk[l] = e.v[l];
  uint32_t w[9];
#pragma unroll
  for (int l = 0; l < 8; ++l) w[l] = __builtin_bswap32(pk[l]);  // big-endian, sign included
  const uint32_t chk = onion_checksum(pk);
  w[8] = ((chk & 0xffu) << 24) | (((chk >> 8) & 0xffu) << 16) | (0x03u << 8);
  const uint32_t r = scan_address(w, banks, f.nbanks);
  const bool walk = !f.hash_addr || window_hit<56, 9>(w, maps, f.hash_ws);
  const bool hit =
      r == 1 ||
      (r == 2 && pattern_slow(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], f.pats,
                              f.npat)) ||
      (f.nregex && walk &&
       regex_address_slow(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], f.regex, f.dfa,
                          f.dfa_end, f.nregex));
  Fe c;
#pragma unroll
  for (int l = 0; l < 8; ++l) c.v[l] = pk[l];
  c.v[7] &= 0x7fffffffu;  // survivors carry the key with the sign clear
  report(hit, (uint64_t)e.v[8] | ((uint64_t)e.v[9] << 32), c, sink);
}

template <uint32_t MODE>
__device__ __forceinline__ void drain(const uint32_t *scratch, uint32_t top, uint32_t T,
                                      uint32_t g, uint32_t n, const DevFilterArgs &f,
                                      const ScanBank *banks, lptr32 maps, const HitSink &sink) {
  for (uint32_t j = 0; j < n; ++j)  // j is uniform over the active lanes
    drain_one(ld10(scratch, top - j, T, g), f, banks, maps, sink);
}

// The sign of x for parked entries that left it out (defer_sign), then the
// address stage of each.  Parking is compacted per wave, so a lane drains
// entries of any chain of its wave: an entry's index k names its chain c
// ((k - k_base) / (2H+1) in a one-round launch), whose window centre P sits
// in planes 0..3 of column c, and the candidate: P - Q_m (k = kc - m - 1) has
// x = (X yq - Y xq) / (Z - T dq), P + Q_m has x = (X yq + Y xq) / (Z + T dq),
// the centre kc has x = X / Z.  The denominators of up to half - 4 entries at
// a time share one inversion (Montgomery's trick), their prefix products held
// in planes 4.. of the lane's own column (the backward pass has consumed every
// plane below the parking planes, which start at plane half).
__device__ __forceinline__ void sign_terms(const Fe10 &e, const uint32_t *scratch, uint32_t T,
                                           uint64_t k_base, uint32_t half, cptr table, Fe10 &num,
                                           Fe10 &den) {
  const uint64_t k = (uint64_t)e.v[8] | ((uint64_t)e.v[9] << 32);
  const uint64_t S = 2ull * half + 1;
  uint64_t c = (k - k_base) / S;
  c = c < T ? c : T - 1u;  // an entry's chain is below T; this only bounds the reads
  const uint64_t kc = k_base + c * S + half;
  Point10 P;
  P.X = ld10(scratch, 0, T, (uint32_t)c);
  P.Y = ld10(scratch, 1, T, (uint32_t)c);
  P.Z = ld10(scratch, 2, T, (uint32_t)c);
  P.T = ld10(scratch, 3, T, (uint32_t)c);
  if (k == kc) {
    num = P.X;
    den = P.Z;
    return;
  }
  const bool minus = k < kc;
  // an entry's index is within half of its centre; the clamp only keeps the
  // table read in bounds
  uint32_t m = (uint32_t)(minus ? kc - k : k - kc) - 1u;
  m = m < half ? m : half - 1u;
  const cptr q = table + m * kTableWords;
  const Fe10 xq = ld10_const(q + kTabX), yq = ld10_const(q + kTabY), dq = ld10_const(q + kTabDq);
  const Fe10 C = fe10_mul(P.T, dq);
  const Fe10 XY = fe10_mul(P.X, yq), YX = fe10_mul(P.Y, xq);
  num = minus ? fe10_sub(XY, YX) : fe10_add(XY, YX);
  den = minus ? fe10_sub(P.Z, C) : fe10_add(P.Z, C);
}

__device__ __forceinline__ void drain_deferred(uint32_t *scratch, uint32_t top, uint32_t half,
                                               uint32_t T, uint32_t g, uint32_t n,
                                               uint64_t k_base, cptr table,
                                               const DevFilterArgs &f, const ScanBank *banks,
                                               lptr32 maps, const HitSink &sink) {
  const uint32_t K = half - 4;  // entries per inversion
  for (uint32_t j0 = 0; j0 < n; j0 += K) {
    const uint32_t cnt = n - j0 < K ? n - j0 : K;
    Fe10 D = fe10_one();
    for (uint32_t i = 0; i < cnt; ++i) {
      st10(scratch, 4 + i, T, g, D);  // the product of the denominators before entry i
      Fe10 num, den;
      sign_terms(ld10(scratch, top - (j0 + i), T, g), scratch, T, k_base, half, table, num, den);
      D = fe10_mul(den, D);
    }
    Fe10 inv = fe10_invert_sg(D);
    for (uint32_t i = cnt; i-- > 0;) {
      Fe10 e = ld10(scratch, top - (j0 + i), T, g);
      Fe10 num, den;
      sign_terms(e, scratch, T, k_base, half, table, num, den);
      const uint32_t sign = fe10_parity_mul(num, fe10_mul(ld10(scratch, 4 + i, T, g), inv));
      inv = fe10_mul(den, inv);
      e.v[7] |= sign << 31;
      drain_one(e, f, banks, maps, sink);
    }
  }
}

// The address stage of one chain's parked candidates, as its own kernel runs
// it: the same drain, with the window's parking planes left by the search.
__device__ __forceinline__ void address_stage(const uint32_t *scratch, uint32_t half, uint32_t T,
                                              uint32_t g, uint32_t n, const DevFilterArgs &f,
                                              const ScanBank *banks, const HitSink &sink) {
  drain<kModePattern>(scratch, kPlanesPerPair * half, T, g, n, f, banks, nullptr, sink);
}

// `rounds` windows of one chain (the method is described in og_kernels.hpp).
// Every multiplication takes its looser operand first (fe10_mul's bounds).  g
// indexes the SoA buffers (stride T); the chain's windows are numbered
// w = r*chains + local from k_base.  `banks` is the automaton in LDS.
template <uint32_t MODE>
__device__ __forceinline__ void search_chain(Point10 &P, const Cached10 &st, cptr table,
                                             uint32_t *scratch, uint32_t T, uint32_t g,
                                             uint32_t H, uint32_t rounds, uint32_t chains,
                                             uint32_t local, uint64_t k_base,
                                             const DevFilterArgs &f, const ScanBank *banks,
                                             const uint16_t *dfa_lds, uint32_t *parked_out,
                                             const HitSink &sink,
                                             const uint32_t *coarse = nullptr) {
  constexpr bool kPat = (MODE & kModePattern) != 0;
  // the sign of x is left to the address kernel (set only for one-round
  // launches with an address kernel: parked_out non-null)
  const bool defer = kPat && f.defer_sign != 0 && parked_out != nullptr;
  const uint64_t S = 2ull * H + 1;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint64_t w = (uint64_t)r * chains + local;
    const uint64_t kc = k_base + w * S + H;
    // the window's products of the centre's coordinates
    const Fe10 YZ = fe10_mul(P.Y, P.Z), XZ = fe10_mul(P.X, P.Z);
    const Fe10 TX = fe10_mul(P.T, P.X), TY = fe10_mul(P.T, P.Y);
    Fe10 D = fe10_one();
    {
      const Fe10 ZZ = fe10_sq(P.Z), TT = fe10_sq(P.T);
      for (uint32_t m = 0; m < H; ++m) {
        const Fe10 E = fe10_add(ZZ, fe10_mulc(TT, ld10_const(table + m * kTableWords + kTabNdq2)));
        st10(scratch, kChainPlanes * m, T, g, D);
        st10(scratch, kChainPlanes * m + 1, T, g, E);
        D = fe10_mul(E, D);
      }
    }
    // inv = 1/(D Z): the centre takes 1/Z = inv D, and inv Z = 1/D starts
    // the backward pass
    Fe10 inv = fe10_invert_sg(fe10_mul(D, P.Z));
    const uint32_t top = kPlanesPerPair * H;  // the parking plane past the window's
    uint32_t parked = 0;  // the wave's parked entries (uniform)
    {
      const Fe10 iz = fe10_mul(inv, D);
      if (kPat) {
        const Fe c = fe10_canon_product(fe10_mul(P.Y, iz));
        uint32_t kw[8];
        key_words(c, kw);
        bool hit, need;
        key_stage<MODE>(c, kw, f, banks, dfa_lds, hit, need);
        const uint64_t mn = __ballot(need);
        if (mn) {
          uint32_t sign = 0;
          if (!defer) sign = fe10_parity_mul(P.X, iz);
          park(scratch, top, T, g, parked, mn, need, c, sign, kc);
        }
        report(hit, kc, c, sink);
      } else {
        test_key_only<MODE>(P.Y, iz, kc, f, sink, coarse);
      }
    }
    inv = fe10_mul(inv, P.Z);
    // Software-pipelined by one pair: the two elements of pair m-1 are loaded
    // as soon as pair m's are consumed, so their latency hides behind the
    // products that form pair m's candidates.
    Fe10 D0 = ld10(scratch, kChainPlanes * (H - 1), T, g);
    Fe10 E0 = ld10(scratch, kChainPlanes * (H - 1) + 1, T, g);
    for (uint32_t m = H; m-- > 0;) {
      // entry: inv = 1/(D0 E0)
      const Fe10 iE = fe10_mul(D0, inv);
      inv = fe10_mul(E0, inv);
      {
        const size_t next = kChainPlanes * (m ? m - 1 : 0);  // pair 0 reloads itself
        D0 = ld10(scratch, next, T, g);
        E0 = ld10(scratch, next + 1, T, g);
      }
      const cptr q = table + m * kTableWords;
      const Fe10 al = fe10_mulc2(YZ, ld10_const(q + kTabY), TX, ld10_const(q + kTabXdq));
      const Fe10 be = fe10_mulc2(XZ, ld10_const(q + kTabX), TY, ld10_const(q + kTabYdq));
      // y(P + Q) = up iE, y(P - Q) = um iE
      const Fe10 up = fe10_add(al, be), um = fe10_sub(al, be);
      const uint64_t km = kc - (m + 1), kp = kc + (m + 1);
      if (kPat) {
        const Fe cm = fe10_canon_product(fe10_mul(um, iE));
        const Fe cp = fe10_canon_product(fe10_mul(up, iE));
        uint32_t wm[8], wp[8];
        key_words(cm, wm);
        key_words(cp, wp);
        bool hm, nm, hp, np;
        key_stage_pair<MODE>(cm, wm, cp, wp, f, banks, dfa_lds, hm, nm, hp, np);
        const uint64_t mm = __ballot(nm), mp = __ballot(np);
        if (defer) {
          if (mm) park(scratch, top, T, g, parked, mm, nm, cm, 0u, km);
          if (mp) park(scratch, top, T, g, parked, mp, np, cp, 0u, kp);
        } else if (mm | mp) {
          // x(P +- Q) = (gamma +- delta) / E, gamma = XZ yq + TY (-xq dq) and
          // delta = YZ xq + TX (-yq dq), each one double product
          const Fe10 ga = fe10_mulc2(XZ, ld10_const(q + kTabY), TY, ld10_const(q + kTabNxdq));
          const Fe10 de = fe10_mulc2(YZ, ld10_const(q + kTabX), TX, ld10_const(q + kTabNydq));
          if (mm)
            park(scratch, top, T, g, parked, mm, nm, cm, fe10_parity_mul(fe10_sub(ga, de), iE),
                 km);
          if (mp)
            park(scratch, top, T, g, parked, mp, np, cp, fe10_parity_mul(fe10_add(ga, de), iE),
                 kp);
        }
        report(hm, km, cm, sink);
        report(hp, kp, cp, sink);
      } else {
        test_key_pair<MODE>(um, up, iE, km, kp, f, sink, coarse);
      }
    }
    if (kPat) {
      // the address stage: here, or (one round per launch) in the address
      // kernel, which runs it at an occupancy the search loop cannot have
      const uint32_t share = parked_share(parked);
      if (parked_out) {
        parked_out[g] = share;
        // the centre, from which the address kernel derives the deferred signs
        if (defer) {
          st10(scratch, 0, T, g, P.X);
          st10(scratch, 1, T, g, P.Y);
          st10(scratch, 2, T, g, P.Z);
          st10(scratch, 3, T, g, P.T);
        }
      } else {
        drain<MODE>(scratch, top, T, g, share, f, banks,
                    f.hash_addr ? (lptr32)dfa_lds : (lptr32) nullptr, sink);
      }
    }
    P = pt10_add1(P, st);
  }
}

// The automaton banks of a filter, and its first expression automaton when
// that fits, into this workgroup's LDS copies.  The tables are whole rows of
// 32 16-bit entries from a 256-byte aligned allocation, so they copy as
// 32-bit words.
// With `with_maps` (the search kernel's LDS holds kPatternLdsWords words) the
// window prefilter's maps are copied instead when the set has them.  Without
// `with_states` the automaton is not copied: the address stage walks the
// expressions' automata in global memory, so its kernels hold only the banks
// (and the maps) in LDS, which lets more workgroups share a CU.
__device__ __forceinline__ void load_filter(ScanBank *lds, uint16_t *dfa_lds,
                                            const DevFilterArgs &f, bool with_maps = false,
                                            bool with_states = true) {
  uint32_t *dst = reinterpret_cast<uint32_t *>(lds);
  for (uint32_t i = threadIdx.x; i < f.nbanks * kBankWords; i += blockDim.x) dst[i] = f.banks[i];
  if (with_maps && f.hash) {
    uint32_t *d = reinterpret_cast<uint32_t *>(dfa_lds);
    for (uint32_t i = threadIdx.x; i < 2 * kHashSliceWords; i += blockDim.x) d[i] = f.hash[i];
  } else if (with_states && f.lds_states) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(f.dfa + f.regex[0]);
    uint32_t *d = reinterpret_cast<uint32_t *>(dfa_lds);
    for (uint32_t i = threadIdx.x; i < f.lds_states * 16; i += blockDim.x) d[i] = src[i];
  }
  __syncthreads();
}

__device__ __forceinline__ Point10 ld_point(const uint32_t *state, uint32_t T, uint32_t g) {
  Point10 P;
  P.X = ld10(state, 0, T, g);
  P.Y = ld10(state, 1, T, g);
  P.Z = ld10(state, 2, T, g);
  P.T = ld10(state, 3, T, g);
  return P;
}

__device__ __forceinline__ void st_point(uint32_t *state, uint32_t T, uint32_t g, const Point10 &P) {
  st10(state, 0, T, g, P.X);
  st10(state, 1, T, g, P.Y);
  st10(state, 2, T, g, P.Z);
  st10(state, 3, T, g, P.T);
}

// Chain start points: P_c = P0 + (k0 + c(2H+1) + H) * 8G, computed in the
// 8 x 32 representation and stored in the search's 10-limb one.
__device__ __forceinline__ void setup_chain(const SetupArgs &a, uint32_t local) {
  Cached eight, start;
  Point p0;
  for (int l = 0; l < 8; ++l) {
    eight.ypx.v[l] = a.eight[l];
    eight.ymx.v[l] = a.eight[8 + l];
    eight.Z.v[l] = a.eight[16 + l];
    eight.t2d.v[l] = a.eight[24 + l];
    p0.X.v[l] = a.start[l];
    p0.Y.v[l] = a.start[8 + l];
    p0.Z.v[l] = a.start[16 + l];
    p0.T.v[l] = a.start[24 + l];
  }
  Fe two_d;
  for (int l = 0; l < 8; ++l) two_d.v[l] = a.two_d[l];
  start = pt_to_cached(p0, two_d);
  const uint64_t S = 2ull * a.half + 1;
  const uint64_t m = a.k0 + (uint64_t)local * S + a.half;
  Point p = pt_add(pt_mul_u64(m, eight), start);
  Point10 q;
  q.X = fe10_from(p.X);
  q.Y = fe10_from(p.Y);
  q.Z = fe10_from(p.Z);
  q.T = fe10_from(p.T);
  st_point(a.state, a.threads, a.first + local, q);
}

}  // namespace og
