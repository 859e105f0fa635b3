// Device side of the search, shared by the single-search engine and the
// batched (slot) engine: SoA access, the filter tests, the wave-level hit
// append and the window loop.  See og_kernels.hpp for the method.
#pragma once
#include "og_kernels.hpp"

namespace og {

typedef const __attribute__((address_space(4))) uint32_t *cptr;

// Occupancy of the search kernels, in waves per SIMD.  Left alone the compiler
// spends about 300 VGPRs on the ten-limb products and runs one wave per SIMD;
// held to two waves (256 VGPRs, a few spills) the kernel measured 20% faster.
#ifndef OG_SEARCH_WAVES
#define OG_SEARCH_WAVES 2
#endif
#if OG_SEARCH_WAVES > 0
#define OG_SEARCH_ATTR __attribute__((amdgpu_waves_per_eu(OG_SEARCH_WAVES, OG_SEARCH_WAVES)))
#else
#define OG_SEARCH_ATTR
#endif

__device__ __forceinline__ Fe10 ld10_const(cptr p) {
  Fe10 r;
#pragma unroll
  for (int l = 0; l < 10; ++l) r.v[l] = p[l];
  return r;
}

// Element `plane` of a structure-of-arrays buffer of 10-limb values: limb l of
// chain g sits at (plane*10 + l)*T + g, so a wave touches 64 consecutive words.
__device__ __forceinline__ Fe10 ld10(const uint32_t *base, size_t plane, uint32_t T, uint32_t g) {
  Fe10 r;
#pragma unroll
  for (int l = 0; l < 10; ++l) r.v[l] = __builtin_nontemporal_load(base + (plane * 10 + l) * T + g);
  return r;
}

__device__ __forceinline__ void st10(uint32_t *base, size_t plane, uint32_t T, uint32_t g,
                                     const Fe10 &v) {
#pragma unroll
  for (int l = 0; l < 10; ++l) __builtin_nontemporal_store(v.v[l], base + (plane * 10 + l) * T + g);
}

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

__device__ __forceinline__ bool bitmap_bit(const DevFilterArgs &f, uint32_t head) {
  const uint32_t slot = (uint32_t)((uint64_t)head >> (32 - f.bits));
  return (f.bitmap[slot >> 5] >> (slot & 31)) & 1u;
}

// The cheap test every candidate gets: on the low 32 bits of the key (its
// first four bytes), for both the value and the value minus p.  Never false
// for a candidate the exact test accepts.
template <uint32_t MODE>
__device__ __forceinline__ bool maybe_key(const Fe10 &y, const DevFilterArgs &f) {
  const uint32_t lo = fe10_low32(y), lo19 = lo + 19u;
  if (MODE & kModePrefix)
    return (((lo ^ f.want[0]) & f.mask[0]) == 0) | (((lo19 ^ f.want[0]) & f.mask[0]) == 0);
  if (MODE & kModeBitmap) return bitmap_bit(f, bswap32(lo)) | bitmap_bit(f, bswap32(lo19));
  return false;
}

// The exact key-side test on the canonical key (sign clear).
template <uint32_t MODE>
__device__ __forceinline__ bool exact_key(const Fe &c, const DevFilterArgs &f) {
  if (MODE & kModePrefix) {
    uint32_t diff = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) diff |= (c.v[l] ^ f.want[l]) & f.mask[l];
    return diff == 0;
  }
  if (MODE & kModeBitmap) {
    const uint32_t head = bswap32(c.v[0]);
    if (!bitmap_bit(f, head)) return false;
    if (f.nbuckets == 0) return true;
    const uint64_t h = ((uint64_t)head << 32) | bswap32(c.v[1]);
    for (uint32_t b = 0; b < f.nbuckets; ++b) {
      const uint32_t off = f.buckets[4 * b], cnt = f.buckets[4 * b + 1];
      const uint64_t mask = (uint64_t)f.buckets[4 * b + 2] | ((uint64_t)f.buckets[4 * b + 3] << 32);
      const uint64_t probe = h & mask;
      const uint64_t *keys = f.keys + off;
      uint32_t lo = 0, hi = cnt;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < probe) lo = mid + 1;
        else hi = mid;
      }
      if (lo < cnt && keys[lo] == probe) return true;
    }
    return false;
  }
  return false;
}

// Symbol m (0..55) of the address held as nine big-endian words.
__device__ __forceinline__ uint32_t addr_symbol(const uint32_t w[9], uint32_t m) {
  const uint32_t bit = 5 * m, i = bit >> 5, sh = bit & 31;
  const uint32_t hi = w[i];
  const uint32_t lo = i + 1 < 9 ? w[i + 1] : 0u;
  const uint64_t both = ((uint64_t)hi << 32) | lo;
  return (uint32_t)(both >> (59 - sh)) & 31u;
}

// The general forms on the full address: pk = canonical y with the sign of x
// in bit 255, then SHA3-256 for the two checksum bytes (src/key.rs:106-132).
__device__ __forceinline__ bool pattern_test(const Fe &cy, const Fe &cx, const DevFilterArgs &f) {
  uint32_t pk[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) pk[l] = cy.v[l];
  pk[7] |= (cx.v[0] & 1u) << 31;
  const uint32_t chk = onion_checksum(pk);
  uint32_t w[9];
#pragma unroll
  for (int l = 0; l < 8; ++l) w[l] = bswap32(pk[l]);
  w[8] = ((chk & 0xffu) << 24) | (((chk >> 8) & 0xffu) << 16) | (0x03u << 8);
  for (uint32_t p = 0; p < f.npat; ++p) {
    const uint32_t *rec = f.pats + p * kPatWords;
    const uint32_t len = rec[0], first = rec[1], last = rec[2];
    for (uint32_t s = first; s <= last; ++s) {
      bool ok = true;
      for (uint32_t i = 0; i < len && ok; ++i) ok = (rec[4 + i] >> addr_symbol(w, s + i)) & 1u;
      if (ok) return true;
    }
  }
  return false;
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
  uint32_t *slot_hits;  // per-slot hit counters, or null
};

template <uint32_t MODE>
__device__ __forceinline__ void test_candidate(const Fe10 &y, const Fe10 &x, uint64_t k,
                                               const DevFilterArgs &f, const HitSink &sink) {
  constexpr bool kKey = (MODE & (kModePrefix | kModeBitmap)) != 0;
  constexpr bool kPat = (MODE & kModePattern) != 0;
  bool hit = false;
  Fe c;
  if (kPat) {
    c = fe10_canon(y);
    if (kKey) hit = exact_key<MODE>(c, f);
    if (!hit) hit = pattern_test(c, fe10_canon(x), f);
  } else {
    const bool maybe = maybe_key<MODE>(y, f);
    if (__ballot(maybe) == 0) return;  // the common case: no lane needs more
    if (maybe) {
      c = fe10_canon(y);
      hit = exact_key<MODE>(c, f);
    }
  }
  if (sink.slot_hits) {
    const uint64_t m = __ballot(hit);
    if (m && __lane_id() == (uint32_t)__ffsll((unsigned long long)m) - 1)
      atomicAdd(sink.slot_hits + sink.slot, (uint32_t)__popcll(m));
  }
  append(hit, sink.slot, k, c, sink.counter, sink.hits, sink.cap);
}

// `rounds` windows of one chain.  Every multiplication takes its looser
// operand first (fe10_mul's bounds).  g indexes the SoA buffers (stride T); the
// chain's windows are numbered w = r*chains + local from k_base.
template <uint32_t MODE>
__device__ __forceinline__ void search_chain(Point10 &P, const Cached10 &st, cptr table,
                                             uint32_t *scratch, uint32_t T, uint32_t g,
                                             uint32_t H, uint32_t rounds, uint32_t chains,
                                             uint32_t local, uint64_t k_base,
                                             const DevFilterArgs &f, const HitSink &sink) {
  constexpr bool kX = (MODE & kModePattern) != 0;
  const uint64_t S = 2ull * H + 1;
  uint32_t *xs = scratch + (size_t)2 * H * 10 * T;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint64_t w = (uint64_t)r * chains + local;
    const uint64_t kc = k_base + w * S + H;
    Fe10 D = fe10_one();
    for (uint32_t m = 0; m < H; ++m) {
      const Fe10 xq = ld10_const(table + m * 30);
      const Fe10 yq = ld10_const(table + m * 30 + 10);
      const Fe10 dq = ld10_const(table + m * 30 + 20);
      const Fe10 A = fe10_mul(P.Y, yq);
      const Fe10 B = fe10_mul(P.X, xq);
      const Fe10 C = fe10_mul(P.T, dq);
      const Fe10 dp = fe10_sub(P.Z, C);
      const Fe10 dm = fe10_add(P.Z, C);
      const Fe10 D0 = D;
      st10(scratch, 2 * m, T, g, fe10_mul(fe10_add(A, B), D));
      D = fe10_mul(dp, D);
      st10(scratch, 2 * m + 1, T, g, fe10_mul(fe10_sub(A, B), D));
      if (kX) {
        const Fe10 XY = fe10_mul(P.X, yq);
        const Fe10 YX = fe10_mul(P.Y, xq);
        st10(xs, 2 * m, T, g, fe10_mul(fe10_add(XY, YX), D));
        st10(xs, 2 * m + 1, T, g, fe10_mul(fe10_sub(XY, YX), D0));
      }
      D = fe10_mul(dm, D);
    }
    const Fe10 sc = fe10_mul(P.Y, D);
    const Fe10 tc = kX ? fe10_mul(P.X, D) : D;
    D = fe10_mul(D, P.Z);
    Fe10 inv = fe10_invert(D);
    {
      const Fe10 y = fe10_mul(sc, inv);
      const Fe10 x = kX ? fe10_mul(tc, inv) : y;
      test_candidate<MODE>(y, x, kc, f, sink);
    }
    inv = fe10_mul(inv, P.Z);
    for (uint32_t m = H; m-- > 0;) {
      const Fe10 dq = ld10_const(table + m * 30 + 20);
      const Fe10 C = fe10_mul(P.T, dq);
      const Fe10 dp = fe10_sub(P.Z, C);
      const Fe10 dm = fe10_add(P.Z, C);
      // entry: inv = 1/(D0 dp dm).  y(P-Q) = s_m inv and x(P+Q) = t_p inv;
      // after inv *= dm: y(P+Q) = s_p inv and x(P-Q) = t_m inv.
      const Fe10 ym = fe10_mul(ld10(scratch, 2 * m + 1, T, g), inv);
      const Fe10 xp = kX ? fe10_mul(ld10(xs, 2 * m, T, g), inv) : ym;
      inv = fe10_mul(dm, inv);
      const Fe10 xm = kX ? fe10_mul(ld10(xs, 2 * m + 1, T, g), inv) : ym;
      test_candidate<MODE>(ym, xm, kc - (m + 1), f, sink);
      const Fe10 yp = fe10_mul(ld10(scratch, 2 * m, T, g), inv);
      test_candidate<MODE>(yp, xp, kc + (m + 1), f, sink);
      inv = fe10_mul(dp, inv);
    }
    P = pt10_add1(P, st);
  }
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

// Chain start points: P_c = P0 + (c(2H+1) + H) * 8G, computed in the 8 x 32
// representation and stored in the search's 10-limb one.
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
  const uint64_t m = (uint64_t)local * S + a.half;
  Point p = pt_add(pt_mul_u64(m, eight), start);
  Point10 q;
  q.X = fe10_from(p.X);
  q.Y = fe10_from(p.Y);
  q.Z = fe10_from(p.Z);
  q.T = fe10_from(p.T);
  st_point(a.state, a.threads, a.first + local, q);
}

}  // namespace og
