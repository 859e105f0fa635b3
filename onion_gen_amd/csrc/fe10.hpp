// The radix-2^25.5 alternative (ten limbs of 26/25 bits, the shape of the
// reference's src/field4.rs:568-670 lane multiply), kept for the radix bench
// that decides which representation the search uses on gfx950.
//
// Limb i holds bits [ceil(25.5 i), ceil(25.5 (i+1))): 26 bits for even i,
// 25 for odd i.  A product needs 100 partial products of at most 32 x 32 bits,
// each on v_mad_u64_u32 with a 64-bit column accumulator; wrapped columns are
// scaled by 19 (2^255 = 19) and odd x odd terms by 2 (the half bit).
#pragma once
#include "fe.hpp"

namespace og {

struct Fe10 {
  uint32_t v[10];
};

OG_HD Fe10 fe10_from(const Fe &a) {
  Fe c = fe_canon(a);
  uint64_t bits[4];
  for (int i = 0; i < 4; ++i) bits[i] = (uint64_t)c.v[2 * i] | ((uint64_t)c.v[2 * i + 1] << 32);
  Fe10 r;
  int pos = 0;
  for (int i = 0; i < 10; ++i) {
    const int w = (i & 1) ? 25 : 26;
    const int word = pos >> 6, sh = pos & 63;
    uint64_t x = bits[word] >> sh;
    if (sh + w > 64 && word + 1 < 4) x |= bits[word + 1] << (64 - sh);
    r.v[i] = (uint32_t)(x & ((1ull << w) - 1));
    pos += w;
  }
  return r;
}

// Carry into 26/25-bit limbs; the carry out of limb 9 re-enters limb 0 as 19.
OG_HD void fe10_carry(uint64_t h[10], Fe10 &r) {
  uint64_t c;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < 10; ++i) {
      const int w = (i & 1) ? 25 : 26;
      c = h[i] >> w;
      h[i] &= (1ull << w) - 1;
      if (i < 9) h[i + 1] += c;
      else h[0] += 19 * c;
    }
  }
  for (int i = 0; i < 10; ++i) r.v[i] = (uint32_t)h[i];
}

OG_HD Fe fe10_to(const Fe10 &a) {
  // Sum the limbs at their bit positions with full carries (limbs may be
  // slightly wider than their nominal width), then fold 2^256 = 38.
  uint32_t w[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int pos = 0;
  for (int i = 0; i < 10; ++i) {
    const int width = (i & 1) ? 25 : 26;
    uint64_t v = a.v[i];
    int word = pos >> 5, sh = pos & 31;
    uint64_t add = v << sh;  // spans at most three 32-bit words
    uint64_t c = 0;
    for (int k = word; k < 10; ++k) {
      c += (uint64_t)w[k] + (uint32_t)add;
      w[k] = (uint32_t)c;
      c >>= 32;
      add >>= 32;
      if (!add && !c) break;
    }
    pos += width;
  }
  uint32_t wide[16] = {0};
  for (int i = 0; i < 10; ++i) wide[i] = w[i];
  return fe_canon(fe_fold(wide));
}

OG_HD Fe10 fe10_mul(const Fe10 &f, const Fe10 &g) {
  uint32_t g19[10], f2[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    g19[i] = 19u * g.v[i];
    f2[i] = (i & 1) ? 2u * f.v[i] : f.v[i];
  }
  uint64_t h[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const int j = k - i;
      if (j >= 0) {
        // odd x odd products carry the extra half bit
        const uint32_t fi = ((i & 1) && (j & 1)) ? f2[i] : f.v[i];
        acc = mad64(fi, g.v[j], acc);
      } else {
        const int jj = j + 10;
        const uint32_t fi = ((i & 1) && (jj & 1)) ? f2[i] : f.v[i];
        acc = mad64(fi, g19[jj], acc);
      }
    }
    h[k] = acc;
  }
  Fe10 r;
  fe10_carry(h, r);
  return r;
}

}  // namespace og
