// Radix 2^51 in five 64-bit limbs, the fourth candidate of the radix bench:
// the form of the reference's fiat-crypto field (src/field.rs:368-432,
// fiat 5 x 51) and of the 64-bit CPU fields generally.
//
// gfx950 has no 64 x 64 -> 128 multiply: each of the 25 limb products is
// four v_mad_u64_u32 of 32-bit halves plus the adds that carry between them,
// and every column sum is 128 bits wide, so the 100 products of the 10 x 25.5
// form come back with carry arithmetic around them.  The bench measures what
// that costs.  Limbs are below 2^52 between operations; 19 * g < 2^57.
#pragma once
#include "fe.hpp"

namespace og {

struct Fe51 {
  uint64_t v[5];
};

constexpr uint64_t kM51 = (1ull << 51) - 1;

OG_HD Fe51 fe51_from(const Fe &a) {
  const Fe c = fe_canon(a);
  uint64_t w[4];
  for (int i = 0; i < 4; ++i) w[i] = (uint64_t)c.v[2 * i] | ((uint64_t)c.v[2 * i + 1] << 32);
  Fe51 r;
  r.v[0] = w[0] & kM51;
  r.v[1] = ((w[0] >> 51) | (w[1] << 13)) & kM51;
  r.v[2] = ((w[1] >> 38) | (w[2] << 26)) & kM51;
  r.v[3] = ((w[2] >> 25) | (w[3] << 39)) & kM51;
  r.v[4] = (w[3] >> 12) & kM51;
  return r;
}

// Carries every limb below 2^51 (the top carry re-enters limb 0 as 19) and
// packs the value below 2^255 into eight words; fe_canon finishes it.
OG_HD Fe fe51_to(const Fe51 &a) {
  uint64_t h[5];
  for (int i = 0; i < 5; ++i) h[i] = a.v[i];
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < 4; ++i) {
      h[i + 1] += h[i] >> 51;
      h[i] &= kM51;
    }
    h[0] += 19 * (h[4] >> 51);
    h[4] &= kM51;
  }
  const uint64_t w0 = h[0] | (h[1] << 51), w1 = (h[1] >> 13) | (h[2] << 38),
                 w2 = (h[2] >> 26) | (h[3] << 25), w3 = (h[3] >> 39) | (h[4] << 12);
  Fe r;
  const uint64_t w[4] = {w0, w1, w2, w3};
  for (int i = 0; i < 4; ++i) {
    r.v[2 * i] = (uint32_t)w[i];
    r.v[2 * i + 1] = (uint32_t)(w[i] >> 32);
  }
  return fe_canon(r);
}

OG_HD Fe51 fe51_mul(const Fe51 &f, const Fe51 &g) {
  typedef unsigned __int128 u128;
  uint64_t g19[5];
  for (int i = 1; i < 5; ++i) g19[i] = 19 * g.v[i];
  u128 h[5];
  h[0] = (u128)f.v[0] * g.v[0] + (u128)f.v[1] * g19[4] + (u128)f.v[2] * g19[3] +
         (u128)f.v[3] * g19[2] + (u128)f.v[4] * g19[1];
  h[1] = (u128)f.v[0] * g.v[1] + (u128)f.v[1] * g.v[0] + (u128)f.v[2] * g19[4] +
         (u128)f.v[3] * g19[3] + (u128)f.v[4] * g19[2];
  h[2] = (u128)f.v[0] * g.v[2] + (u128)f.v[1] * g.v[1] + (u128)f.v[2] * g.v[0] +
         (u128)f.v[3] * g19[4] + (u128)f.v[4] * g19[3];
  h[3] = (u128)f.v[0] * g.v[3] + (u128)f.v[1] * g.v[2] + (u128)f.v[2] * g.v[1] +
         (u128)f.v[3] * g.v[0] + (u128)f.v[4] * g19[4];
  h[4] = (u128)f.v[0] * g.v[4] + (u128)f.v[1] * g.v[3] + (u128)f.v[2] * g.v[2] +
         (u128)f.v[3] * g.v[1] + (u128)f.v[4] * g.v[0];
  Fe51 r;
  uint64_t c;
  r.v[0] = (uint64_t)h[0] & kM51;
  h[1] += (uint64_t)(h[0] >> 51);
  r.v[1] = (uint64_t)h[1] & kM51;
  h[2] += (uint64_t)(h[1] >> 51);
  r.v[2] = (uint64_t)h[2] & kM51;
  h[3] += (uint64_t)(h[2] >> 51);
  r.v[3] = (uint64_t)h[3] & kM51;
  h[4] += (uint64_t)(h[3] >> 51);
  r.v[4] = (uint64_t)h[4] & kM51;
  c = (uint64_t)(h[4] >> 51);
  r.v[0] += 19 * c;
  r.v[1] += r.v[0] >> 51;
  r.v[0] &= kM51;
  return r;
}

}  // namespace og
