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
// The plane base is wave-uniform and the in-plane index fits 32 bits, so each
// access is a scalar base plus a 32-bit lane offset (no 64-bit address math
// per load).
__device__ __forceinline__ const uint32_t *soa_at(const uint32_t *base, size_t plane, uint32_t T,
                                                  uint32_t g, int l) {
  const char *pl = reinterpret_cast<const char *>(base + plane * 10 * T);
  return reinterpret_cast<const uint32_t *>(pl + (uint32_t)((l * T + g) * 4u));
}

__device__ __forceinline__ Fe10 ld10(const uint32_t *base, size_t plane, uint32_t T, uint32_t g) {
  Fe10 r;
#pragma unroll
  for (int l = 0; l < 10; ++l) r.v[l] = __builtin_nontemporal_load(soa_at(base, plane, T, g, l));
  return r;
}

__device__ __forceinline__ void st10(uint32_t *base, size_t plane, uint32_t T, uint32_t g,
                                     const Fe10 &v) {
#pragma unroll
  for (int l = 0; l < 10; ++l)
    __builtin_nontemporal_store(v.v[l], const_cast<uint32_t *>(soa_at(base, plane, T, g, l)));
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

// The general forms symbol by symbol on the full address: the exact check for
// patterns longer than one automaton bank.
__device__ __forceinline__ bool pattern_test(const uint32_t w[9], const DevFilterArgs &f) {
  for (uint32_t p = 0; p < f.npat; ++p) {
    const uint32_t *rec = f.pats + p * kPatWords;
    const uint32_t len = rec[0], first = rec[1], last = rec[2];
    for (uint32_t s = first; s <= last; ++s) {
      bool ok = true;
      for (uint32_t i = 0; i < len && ok; ++i) ok = (rec[4 + i] >> sym_at(w, (int)(s + i))) & 1u;
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

__device__ __forceinline__ void report(bool hit, uint64_t k, const Fe &c, const HitSink &sink) {
  if (sink.slot_hits) {
    const uint64_t m = __ballot(hit);
    if (m && __lane_id() == (uint32_t)__ffsll((unsigned long long)m) - 1)
      atomicAdd(sink.slot_hits + sink.slot, (uint32_t)__popcll(m));
  }
  append(hit, sink.slot, k, c, sink.counter, sink.hits, sink.cap);
}

// Key forms only: the low-32-bit probe, and the exact test on the lanes the
// wave ballot names.
template <uint32_t MODE>
__device__ __forceinline__ void test_key_only(const Fe10 &y, uint64_t k, const DevFilterArgs &f,
                                              const HitSink &sink) {
  const bool maybe = maybe_key<MODE>(y, f);
  if (__ballot(maybe) == 0) return;  // the common case: no lane needs more
  bool hit = false;
  Fe c;
  if (maybe) {
    c = fe10_canon(y);
    hit = exact_key<MODE>(c, f);
  }
  report(hit, k, c, sink);
}

// The key-derived part of a candidate under the general forms (and the key
// forms when the set has both): `hit` is final, `need` asks for the address.
template <uint32_t MODE>
__device__ __forceinline__ void key_stage(const Fe &c, const uint32_t w[8], const DevFilterArgs &f,
                                          const ScanBank *banks, bool &hit, bool &need) {
  scan_key(w, banks, f.nbanks, hit, need);
  if (MODE & (kModePrefix | kModeBitmap)) hit |= exact_key<MODE>(c, f);
  need &= !hit;
}

// The address stage for the lanes that need it: the sign of x (bit 0 of its
// canonical form), SHA3-256 for the checksum, the exact automaton pass.
__device__ __forceinline__ bool address_stage(const uint32_t kw[8], const Fe10 &x,
                                              const DevFilterArgs &f, const ScanBank *banks) {
  const uint32_t sign = fe10_canon(x).v[0] & 1u;
  uint32_t pk[8];
#pragma unroll
  for (int l = 0; l < 7; ++l) pk[l] = __builtin_bswap32(kw[l]);
  pk[7] = __builtin_bswap32(kw[7]) | (sign << 31);
  uint32_t w[9];
  address_words(kw, sign, onion_checksum(pk), w);
  const uint32_t r = scan_address(w, banks, f.nbanks);
  return r == 1 || (r == 2 && pattern_test(w, f));
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
                                             const HitSink &sink) {
  constexpr bool kPat = (MODE & kModePattern) != 0;
  const uint64_t S = 2ull * H + 1;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint64_t w = (uint64_t)r * chains + local;
    const uint64_t kc = k_base + w * S + H;
    Fe10 D = fe10_one();
    for (uint32_t m = 0; m < H; ++m) {
      const Fe10 dq = ld10_const(table + m * 30 + 20);
      const Fe10 C = fe10_mul(P.T, dq);
      const Fe10 dp = fe10_sub(P.Z, C);
      const Fe10 dm = fe10_add(P.Z, C);
      st10(scratch, kPlanesPerPair * m, T, g, D);
      D = fe10_mul(dp, D);
      st10(scratch, kPlanesPerPair * m + 1, T, g, D);
      st10(scratch, kPlanesPerPair * m + 2, T, g, C);
      D = fe10_mul(dm, D);
    }
    // inv = 1/(D Z): the centre takes 1/Z = inv D, and inv Z = 1/D starts
    // the backward pass
    Fe10 inv = fe10_invert(fe10_mul(D, P.Z));
    {
      const Fe10 iz = fe10_mul(inv, D);
      const Fe10 y = fe10_mul(P.Y, iz);
      if (kPat) {
        const Fe c = fe10_canon(y);
        uint32_t kw[8];
        key_words(c, kw);
        bool hit, need;
        key_stage<MODE>(c, kw, f, banks, hit, need);
        if (__ballot(need)) {
          const Fe10 x = fe10_mul(P.X, iz);
          if (need) hit = address_stage(kw, x, f, banks);
        }
        report(hit, kc, c, sink);
      } else {
        test_key_only<MODE>(y, kc, f, sink);
      }
    }
    inv = fe10_mul(inv, P.Z);
    // Software-pipelined by one pair: the three elements of pair m-1 are
    // loaded as soon as pair m's are consumed, so their latency hides behind
    // the four multiplications that form pair m's candidates.
    Fe10 D0 = ld10(scratch, kPlanesPerPair * (H - 1), T, g);
    Fe10 D1 = ld10(scratch, kPlanesPerPair * (H - 1) + 1, T, g);
    Fe10 C = ld10(scratch, kPlanesPerPair * (H - 1) + 2, T, g);
    for (uint32_t m = H; m-- > 0;) {
      const Fe10 dp = fe10_sub(P.Z, C);
      const Fe10 dm = fe10_add(P.Z, C);
      // entry: inv = 1/(D0 dp dm)
      const Fe10 idm = fe10_mul(D1, inv);
      inv = fe10_mul(dm, inv);
      const Fe10 idp = fe10_mul(D0, inv);
      inv = fe10_mul(dp, inv);
      {
        const size_t next = kPlanesPerPair * (m ? m - 1 : 0);  // pair 0 reloads itself
        D0 = ld10(scratch, next, T, g);
        D1 = ld10(scratch, next + 1, T, g);
        C = ld10(scratch, next + 2, T, g);
      }
      const Fe10 xq = ld10_const(table + m * 30);
      const Fe10 yq = ld10_const(table + m * 30 + 10);
      const Fe10 A = fe10_mul(P.Y, yq);
      const Fe10 B = fe10_mul(P.X, xq);
      const Fe10 ym = fe10_mul(fe10_sub(A, B), idm);  // y(P - Q)
      const Fe10 yp = fe10_mul(fe10_add(A, B), idp);  // y(P + Q)
      const uint64_t km = kc - (m + 1), kp = kc + (m + 1);
      if (kPat) {
        const Fe cm = fe10_canon(ym), cp = fe10_canon(yp);
        uint32_t wm[8], wp[8];
        key_words(cm, wm);
        key_words(cp, wp);
        bool hm, nm, hp, np;
        key_stage<MODE>(cm, wm, f, banks, hm, nm);
        key_stage<MODE>(cp, wp, f, banks, hp, np);
        if (__ballot(nm | np)) {
          // x(P - Q) has the denominator dp, x(P + Q) the denominator dm
          const Fe10 XY = fe10_mul(P.X, yq);
          const Fe10 YX = fe10_mul(P.Y, xq);
          if (nm) hm = address_stage(wm, fe10_mul(fe10_sub(XY, YX), idp), f, banks);
          if (np) hp = address_stage(wp, fe10_mul(fe10_add(XY, YX), idm), f, banks);
        }
        report(hm, km, cm, sink);
        report(hp, kp, cp, sink);
      } else {
        test_key_only<MODE>(ym, km, f, sink);
        test_key_only<MODE>(yp, kp, f, sink);
      }
    }
    P = pt10_add1(P, st);
  }
}

// The automaton banks of a filter into this workgroup's LDS copy.
__device__ __forceinline__ void load_banks(ScanBank *lds, const DevFilterArgs &f) {
  uint32_t *dst = reinterpret_cast<uint32_t *>(lds);
  for (uint32_t i = threadIdx.x; i < f.nbanks * kBankWords; i += blockDim.x) dst[i] = f.banks[i];
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
