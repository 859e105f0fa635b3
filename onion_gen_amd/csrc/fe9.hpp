// Radix 2^29 in nine limbs (261 bits), a third candidate of the radix bench.
//
// With limbs below 2^30, a column of the 9 x 9 schoolbook holds at most nine
// products below 2^60, so every column accumulates in one 64-bit register
// with v_mad_u64_u32 alone: no carry flags in the product.  The carries are
// resolved once per column afterwards, and 2^261 = 2^6 * 2^255 = 64 * 19 = 1216
// (mod p) folds the upper half back.
#pragma once
#include "fe.hpp"

namespace og {

struct Fe9 {
  uint32_t v[9];
};

constexpr uint32_t kM29 = (1u << 29) - 1;

OG_HD Fe9 fe9_from(const Fe &a) {
  Fe c = fe_canon(a);
  Fe9 r;
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, w = bit >> 5, sh = bit & 31;
    uint64_t x = c.v[w] >> sh;
    if (w + 1 < 8) x |= (uint64_t)c.v[w + 1] << (32 - sh);
    r.v[i] = (uint32_t)x & kM29;
  }
  return r;
}

OG_HD Fe fe9_to(const Fe9 &a) {
  uint32_t w[16] = {0};
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, k = bit >> 5, sh = bit & 31;
    uint64_t add = (uint64_t)a.v[i] << sh;
    uint64_t c = 0;
    for (int j = k; j < 16; ++j) {
      c += (uint64_t)w[j] + (uint32_t)add;
      w[j] = (uint32_t)c;
      c >>= 32;
      add >>= 32;
      if (!add && !c) break;
    }
  }
  return fe_canon(fe_fold(w));
}

OG_HD Fe9 fe9_mul(const Fe9 &a, const Fe9 &b) {
  uint64_t col[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int j = k - i;
      if (j >= 0 && j < 9) acc = mad64(a.v[i], b.v[j], acc);
    }
    col[k] = acc;
  }
  uint32_t d[18];
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    const uint64_t t = col[k] + carry;
    d[k] = (uint32_t)t & kM29;
    carry = t >> 29;
  }
  d[17] = (uint32_t)carry;
  Fe9 r;
  carry = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const uint64_t t = mad64(d[k + 9], 1216u, (uint64_t)d[k] + carry);
    r.v[k] = (uint32_t)t & kM29;
    carry = t >> 29;
  }
  r.v[0] += (uint32_t)carry * 1216u;  // below 2^30: a loose limb
  return r;
}

}  // namespace og
