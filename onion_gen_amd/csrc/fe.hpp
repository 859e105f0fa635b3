// GF(2^255 - 19) in eight saturated 32-bit limbs, least significant first.
//
// A value is any 256-bit word congruent to the element (below 2^256, not
// necessarily below p), folded with 2^256 = 38 (mod p) after every operation:
// the representation of the reference's src/field.rs:1-30 (4 x 64) and of its
// device kernel (cuda-kernel/src/field.rs:1-10, 8 x 32), written for gfx950's
// v_mad_u64_u32 (32x32+64 -> 64) and 64-bit adds.  The same code compiles for
// the host, where the library uses it for set-up and hit reconstruction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OG_HD __host__ __device__ __forceinline__

namespace og {

struct Fe {
  uint32_t v[8];
};

OG_HD uint64_t mad64(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

OG_HD Fe fe_zero() {
  Fe r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r.v[i] = 0;
  return r;
}

OG_HD Fe fe_one() {
  Fe r = fe_zero();
  r.v[0] = 1;
  return r;
}

// Fold a 512-bit product w[0..15] to 256 bits: low + 38 * high, then the carry
// out of the top (below 40) comes back as 38 * carry.  A second carry can only
// follow a wrap that left the low limbs near zero, so it is added to limb 0
// without further propagation (src/field.rs:245-268 does both passes in full).
OG_HD Fe fe_fold(const uint32_t w[16]) {
  Fe r;
  uint64_t t = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    t = mad64(w[k + 8], 38u, (uint64_t)w[k] + (t >> 32));
    r.v[k] = (uint32_t)t;
  }
  uint64_t c = (t >> 32) * 38u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += r.v[k];
    r.v[k] = (uint32_t)c;
    c >>= 32;
  }
  r.v[0] += (uint32_t)c * 38u;
  return r;
}

// Schoolbook 8 x 8, one row per limb of a: each partial product plus the
// running column word plus the row carry fits 64 bits exactly
// ((2^32-1)^2 + 2(2^32-1) = 2^64 - 1).
OG_HD Fe fe_mul(const Fe &a, const Fe &b) {
  uint32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint64_t t = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      t = mad64(a.v[i], b.v[j], (uint64_t)w[i + j] + (t >> 32));
      w[i + j] = (uint32_t)t;
    }
    w[i + 8] = (uint32_t)(t >> 32);
  }
  return fe_fold(w);
}

// Squaring: the 28 cross products once, doubled, plus the 8 squares.
OG_HD Fe fe_sq(const Fe &a) {
  uint32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    uint64_t t = 0;
#pragma unroll
    for (int j = i + 1; j < 8; ++j) {
      t = mad64(a.v[i], a.v[j], (uint64_t)w[i + j] + (t >> 32));
      w[i + j] = (uint32_t)t;
    }
    w[i + 8] = (uint32_t)(t >> 32);
  }
  // double the cross products, then add the diagonal squares
  uint32_t top = 0;
#pragma unroll
  for (int k = 15; k > 0; --k) w[k] = (w[k] << 1) | (w[k - 1] >> 31);
  w[0] <<= 1;
  (void)top;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint64_t s = (uint64_t)a.v[i] * a.v[i];
    c += (uint64_t)w[2 * i] + (uint32_t)s;
    w[2 * i] = (uint32_t)c;
    c >>= 32;
    c += (uint64_t)w[2 * i + 1] + (uint32_t)(s >> 32);
    w[2 * i + 1] = (uint32_t)c;
    c >>= 32;
  }
  return fe_fold(w);
}

// a + b; the carry out of the top returns as 38 (src/field.rs:130-144).
OG_HD Fe fe_add(const Fe &a, const Fe &b) {
  Fe r;
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += (uint64_t)a.v[k] + b.v[k];
    r.v[k] = (uint32_t)c;
    c >>= 32;
  }
  c *= 38u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += r.v[k];
    r.v[k] = (uint32_t)c;
    c >>= 32;
  }
  r.v[0] += (uint32_t)c * 38u;
  return r;
}

// a - b; a borrow out of the top means the word wrapped by 2^256 = 38, so 38
// is taken off again (src/field.rs:147-160).  A second borrow leaves limb 0
// near 2^32, which absorbs the last 38 without a further borrow.
OG_HD Fe fe_sub(const Fe &a, const Fe &b) {
  Fe r;
  int64_t c = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += (int64_t)a.v[k] - (int64_t)b.v[k];
    r.v[k] = (uint32_t)c;
    c >>= 32;  // 0 or -1
  }
  c *= 38;  // 0 or -38
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += (int64_t)r.v[k];
    r.v[k] = (uint32_t)c;
    c >>= 32;
  }
  r.v[0] += (uint32_t)(c * 38);  // c is 0 or -1
  return r;
}

// The canonical representative below p: fold bit 255 in as 19, then subtract p
// when x + 19 reaches 2^255.  Equivalent to the three trial subtractions of
// src/field.rs:46-69 for every input below 2^256.
OG_HD Fe fe_canon(const Fe &a) {
  Fe x = a;
  uint32_t top = x.v[7] >> 31;
  x.v[7] &= 0x7fffffffu;
  uint64_t c = 19u * top;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += x.v[k];
    x.v[k] = (uint32_t)c;
    c >>= 32;
  }
  Fe t;
  c = 19;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += x.v[k];
    t.v[k] = (uint32_t)c;
    c >>= 32;
  }
  uint32_t ge = t.v[7] >> 31;  // x >= p
  uint32_t m = 0u - ge;
  t.v[7] &= 0x7fffffffu;
#pragma unroll
  for (int k = 0; k < 8; ++k) x.v[k] = (t.v[k] & m) | (x.v[k] & ~m);
  return x;
}

OG_HD Fe fe_sqn(Fe a, int n) {
  for (int i = 0; i < n; ++i) a = fe_sq(a);
  return a;
}

// x^(p-2) by the ref10 chain: 254 squarings and 11 multiplications
// (src/field.rs:163-220).
OG_HD Fe fe_invert(const Fe &z1) {
  Fe z2 = fe_sq(z1);
  Fe z8 = fe_sqn(z2, 2);
  Fe z9 = fe_mul(z1, z8);
  Fe z11 = fe_mul(z2, z9);
  Fe z22 = fe_sq(z11);
  Fe z5 = fe_mul(z9, z22);
  Fe z10 = fe_mul(fe_sqn(z5, 5), z5);
  Fe z20 = fe_mul(fe_sqn(z10, 10), z10);
  Fe z40 = fe_mul(fe_sqn(z20, 20), z20);
  Fe z50 = fe_mul(fe_sqn(z40, 10), z10);
  Fe z100 = fe_mul(fe_sqn(z50, 50), z50);
  Fe z200 = fe_mul(fe_sqn(z100, 100), z100);
  Fe t = fe_mul(fe_sqn(z200, 50), z50);
  return fe_mul(fe_sqn(t, 5), z11);
}

OG_HD bool fe_is_zero(const Fe &a) {
  Fe c = fe_canon(a);
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) o |= c.v[k];
  return o == 0;
}

OG_HD Fe fe_from_bytes(const uint8_t b[32]) {
  Fe r;
  for (int k = 0; k < 8; ++k)
    r.v[k] = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) | ((uint32_t)b[4 * k + 2] << 16) |
             ((uint32_t)b[4 * k + 3] << 24);
  return r;
}

OG_HD void fe_to_bytes(uint8_t out[32], const Fe &a) {
  Fe c = fe_canon(a);
  for (int k = 0; k < 8; ++k)
    for (int i = 0; i < 4; ++i) out[4 * k + i] = (uint8_t)(c.v[k] >> (8 * i));
}

}  // namespace og
