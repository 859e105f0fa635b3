// GF(2^255 - 19) in radix 2^25.5: ten limbs of 26 / 25 bits (even / odd
// index), the representation the device search runs in.
//
// Why this radix on gfx950 (see the radix bench): v_mad_u64_u32 issues at
// nearly full rate, so the 100 partial products of a 10 x 10 schoolbook cost
// about what the 64 of an 8 x 32 product do, but here every column
// accumulates in one 64-bit register with no carry flag (gfx950 needs two wait
// states between a VALU carry write and its read), and the six or seven free
// bits per limb let additions and subtractions skip carry chains entirely.
// It is the shape of the reference's src/field4.rs:568-670 lane multiply.
//
// Limb i holds bits [a_i, a_{i+1}) with a_i = ceil(25.5 i).  Bounds, in units
// of the limb's nominal width 2^w (w = 26 even, 25 odd):
//   tight  : the result of fe10_mul / fe10_carry, below 1 (limb 0 below 1 + 2^-21)
//   loose  : the sum of two tight values (< 2), a tight minus a tight (< 3), or
//            a loose-2 minus a tight (< 4)
// fe10_mul(f, g) accepts f limbs up to 5 units (its largest column then stays
// below 2^63.6) and g limbs below 3.3 units: g is scaled by 19 in 32 bits
// (19 * 3.3 * 2^26 < 2^32), so the looser operand always goes first.
#pragma once
#include "fe.hpp"

namespace og {

struct Fe10 {
  uint32_t v[10];
};

OG_HD constexpr int fe10_width(int i) { return (i & 1) ? 25 : 26; }

OG_HD Fe10 fe10_zero() {
  Fe10 r;
#pragma unroll
  for (int i = 0; i < 10; ++i) r.v[i] = 0;
  return r;
}

OG_HD Fe10 fe10_one() {
  Fe10 r = fe10_zero();
  r.v[0] = 1;
  return r;
}

// Carry the 64-bit column sums of a product into tight limbs; the carry out of
// limb 9 re-enters limb 0 as 19 (2^255 = 19).  A second, short pass moves the
// (at most 2^38) re-entered carry on by one limb, which leaves limb 0 below
// 2^26 + 19*2^13 and every other limb below 2^w.
OG_HD Fe10 fe10_carry64(uint64_t h[10]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int w = fe10_width(i);
    h[i + 1] += h[i] >> w;
    h[i] &= (1ull << w) - 1;
  }
  const uint64_t c9 = h[9] >> 25;
  h[9] &= (1ull << 25) - 1;
  h[0] += 19 * c9;
  h[1] += h[0] >> 26;
  h[0] &= (1ull << 26) - 1;
  Fe10 r;
#pragma unroll
  for (int i = 0; i < 10; ++i) r.v[i] = (uint32_t)h[i];
  return r;
}

// The scaled limb of a wave-uniform constant g that multiplies limb i of f in
// column k of a product: odd-by-odd terms land half a bit high (x2), and terms
// that wrap past limb 9 re-enter as 19 (2^255 = 19).  38 * 2^25 < 2^31, so
// every scaled limb fits 32 bits.
OG_HD uint32_t fe10_cterm(const Fe10 &g, int k, int i) {
  const int j = k - i;
  const int jj = j < 0 ? j + 10 : j;
  const uint32_t scale = ((i & 1) && (jj & 1) ? 2u : 1u) * (j < 0 ? 19u : 1u);
  return scale * g.v[jj];
}

#ifndef __HIP_DEVICE_COMPILE__
// The product by a tight, wave-uniform g (a table constant): the same
// partial products as fe10_mul, with the doubling of odd-by-odd terms moved
// onto g, so that every scaled form of g (2g, 19g, 38g) is scalar work and
// v_mad_u64_u32 reads it from an SGPR; f needs no preparation at all.
OG_HD Fe10 fe10_mulc(const Fe10 &f, const Fe10 &g) {
  uint64_t h[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) acc = mad64(f.v[i], fe10_cterm(g, k, i), acc);
    h[k] = acc;
  }
  return fe10_carry64(h);
}

// f1*g1 + f2*g2 for tight f1, f2 and wave-uniform tight g1, g2, with one
// carry pass: each column sums twenty products (below 2^61.6).
OG_HD Fe10 fe10_mulc2(const Fe10 &f1, const Fe10 &g1, const Fe10 &f2, const Fe10 &g2) {
  uint64_t h[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) acc = mad64(f1.v[i], fe10_cterm(g1, k, i), acc);
#pragma unroll
    for (int i = 0; i < 10; ++i) acc = mad64(f2.v[i], fe10_cterm(g2, k, i), acc);
    h[k] = acc;
  }
  return fe10_carry64(h);
}
#endif

#ifdef __HIP_DEVICE_COMPILE__
// The device product runs column by column, each column's ten partial
// products accumulated on top of the carry out of the column before, so that
// the carry costs no 64-bit add of its own and only one column sum is live at
// a time (the key forms' search loop then fits three waves per SIMD).  Written
// in plain C the compiler reassociates every column sum and adds the incoming
// carry last; one asm block per column keeps the order: ten dependent
// v_mad_u64_u32 (the compiler emits such dependent sequences back to back
// itself, with no wait states).  The carry-out SGPR pair of each is unused.
__device__ __forceinline__ uint64_t fe10_column(const uint32_t a[10], const uint32_t b[10],
                                                uint64_t c) {
  uint64_t r, cc;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %22\n\t"
      "v_mad_u64_u32 %0, %1, %4, %5, %0\n\t"
      "v_mad_u64_u32 %0, %1, %6, %7, %0\n\t"
      "v_mad_u64_u32 %0, %1, %8, %9, %0\n\t"
      "v_mad_u64_u32 %0, %1, %10, %11, %0\n\t"
      "v_mad_u64_u32 %0, %1, %12, %13, %0\n\t"
      "v_mad_u64_u32 %0, %1, %14, %15, %0\n\t"
      "v_mad_u64_u32 %0, %1, %16, %17, %0\n\t"
      "v_mad_u64_u32 %0, %1, %18, %19, %0\n\t"
      "v_mad_u64_u32 %0, %1, %20, %21, %0"
      : "=&v"(r), "=&s"(cc)
      : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]),
        "v"(a[4]), "v"(b[4]), "v"(a[5]), "v"(b[5]), "v"(a[6]), "v"(b[6]), "v"(a[7]), "v"(b[7]),
        "v"(a[8]), "v"(b[8]), "v"(a[9]), "v"(b[9]), "v"(c));
  return r;
}

// A column summed from zero (the addend of the first product is the inline
// constant 0).
__device__ __forceinline__ uint64_t fe10_column0(const uint32_t a[10], const uint32_t b[10]) {
  uint64_t r, cc;
  asm("v_mad_u64_u32 %0, %1, %2, %3, 0\n\t"
      "v_mad_u64_u32 %0, %1, %4, %5, %0\n\t"
      "v_mad_u64_u32 %0, %1, %6, %7, %0\n\t"
      "v_mad_u64_u32 %0, %1, %8, %9, %0\n\t"
      "v_mad_u64_u32 %0, %1, %10, %11, %0\n\t"
      "v_mad_u64_u32 %0, %1, %12, %13, %0\n\t"
      "v_mad_u64_u32 %0, %1, %14, %15, %0\n\t"
      "v_mad_u64_u32 %0, %1, %16, %17, %0\n\t"
      "v_mad_u64_u32 %0, %1, %18, %19, %0\n\t"
      "v_mad_u64_u32 %0, %1, %20, %21, %0"
      : "=&v"(r), "=&s"(cc)
      : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]),
        "v"(a[4]), "v"(b[4]), "v"(a[5]), "v"(b[5]), "v"(a[6]), "v"(b[6]), "v"(a[7]), "v"(b[7]),
        "v"(a[8]), "v"(b[8]), "v"(a[9]), "v"(b[9]));
  return r;
}

// The same column with the b operands in SGPRs: the scaled limbs of a
// wave-uniform constant (v_mad_u64_u32 reads one SGPR per instruction).
__device__ __forceinline__ uint64_t fe10_column_s(const uint32_t a[10], const uint32_t b[10],
                                                  uint64_t c) {
  uint64_t r, cc;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %22\n\t"
      "v_mad_u64_u32 %0, %1, %4, %5, %0\n\t"
      "v_mad_u64_u32 %0, %1, %6, %7, %0\n\t"
      "v_mad_u64_u32 %0, %1, %8, %9, %0\n\t"
      "v_mad_u64_u32 %0, %1, %10, %11, %0\n\t"
      "v_mad_u64_u32 %0, %1, %12, %13, %0\n\t"
      "v_mad_u64_u32 %0, %1, %14, %15, %0\n\t"
      "v_mad_u64_u32 %0, %1, %16, %17, %0\n\t"
      "v_mad_u64_u32 %0, %1, %18, %19, %0\n\t"
      "v_mad_u64_u32 %0, %1, %20, %21, %0"
      : "=&v"(r), "=&s"(cc)
      : "v"(a[0]), "s"(b[0]), "v"(a[1]), "s"(b[1]), "v"(a[2]), "s"(b[2]), "v"(a[3]), "s"(b[3]),
        "v"(a[4]), "s"(b[4]), "v"(a[5]), "s"(b[5]), "v"(a[6]), "s"(b[6]), "v"(a[7]), "s"(b[7]),
        "v"(a[8]), "s"(b[8]), "v"(a[9]), "s"(b[9]), "v"(c));
  return r;
}

// The column carries of a product summed column by column: limb k keeps the
// low w bits of the column sum, the rest rides into column k + 1, and the
// carry out of limb 9 re-enters limb 0 as 19 and moves on by one limb.
__device__ __forceinline__ void fe10_take(Fe10 &r, int k, uint64_t &acc) {
  const int w = fe10_width(k);
  r.v[k] = (uint32_t)acc & ((1u << w) - 1);
  acc >>= w;
}

__device__ __forceinline__ void fe10_fold(Fe10 &r, uint64_t acc) {
  const uint64_t h0 = (uint64_t)r.v[0] + 19 * acc;
  r.v[1] += (uint32_t)(h0 >> 26);
  r.v[0] = (uint32_t)h0 & ((1u << 26) - 1);
}

// The same fold for a carry below 2^32, one v_mad_u64_u32 where the 64-bit
// carry takes two more instructions.  Column 9 has no wrapped and no doubled
// terms, so its sum is below 10 F G 2^51 (plus a carry in below 2^38) for
// operands below F and G units, and the carry out of limb 9 (that sum over
// 2^25) is below 2^32 whenever F G < 6.4: every product by a constant (F <= 5,
// G = 1, or two of them with F = G = 1) and every multiplication of the
// search loops (F G <= 3).  The curve addition's products (up to 4 x 3
// units) take fe10_mul_wide.
__device__ __forceinline__ void fe10_fold32(Fe10 &r, uint64_t acc) {
  const uint64_t h0 = mad64((uint32_t)acc, 19u, (uint64_t)r.v[0]);
  r.v[1] += (uint32_t)(h0 >> 26);
  r.v[0] = (uint32_t)h0 & ((1u << 26) - 1);
}

// The product by a tight, wave-uniform g (a table constant): the same
// partial products as fe10_mul, with the doubling of odd-by-odd terms moved
// onto g, so that every scaled form of g (2g, 19g, 38g) is scalar work and
// v_mad_u64_u32 reads it from an SGPR; f needs no preparation at all.  The
// columns are summed on the incoming carry, as in fe10_mul.
__device__ __forceinline__ Fe10 fe10_mulc(const Fe10 &f, const Fe10 &g) {
  Fe10 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint32_t b[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) b[i] = fe10_cterm(g, k, i);
    if (k == 0) {
#pragma unroll
      for (int i = 0; i < 10; ++i) acc = mad64(f.v[i], b[i], acc);
    } else {
      acc = fe10_column_s(f.v, b, acc);
    }
    fe10_take(r, k, acc);
  }
  fe10_fold32(r, acc);
  return r;
}

// f1*g1 + f2*g2 for tight f1, f2 and wave-uniform tight g1, g2, with one
// carry pass: each column sums twenty products (below 2^61.6).
__device__ __forceinline__ Fe10 fe10_mulc2(const Fe10 &f1, const Fe10 &g1, const Fe10 &f2,
                                           const Fe10 &g2) {
  Fe10 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint32_t b1[10], b2[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      b1[i] = fe10_cterm(g1, k, i);
      b2[i] = fe10_cterm(g2, k, i);
    }
    if (k == 0) {
#pragma unroll
      for (int i = 0; i < 10; ++i) acc = mad64(f1.v[i], b1[i], acc);
    } else {
      acc = fe10_column_s(f1.v, b1, acc);
    }
    acc = fe10_column_s(f2.v, b2, acc);
    fe10_take(r, k, acc);
  }
  fe10_fold32(r, acc);
  return r;
}

template <bool kWide>
__device__ __forceinline__ Fe10 fe10_mul_t(const Fe10 &f, const Fe10 &g) {
  uint32_t g19[10], f2[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    g19[i] = 19u * g.v[i];
    f2[i] = (i & 1) ? 2u * f.v[i] : f.v[i];
  }
  Fe10 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint32_t a[10], b[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const int j = k - i;
      const int jj = j < 0 ? j + 10 : j;
      a[i] = ((i & 1) && (jj & 1)) ? f2[i] : f.v[i];
      b[i] = j < 0 ? g19[jj] : g.v[j];
    }
    if (k == 0) {  // no carry in: any order of the sum will do
#pragma unroll
      for (int i = 0; i < 10; ++i) acc = mad64(a[i], b[i], acc);
    } else {
      acc = fe10_column(a, b, acc);
    }
    fe10_take(r, k, acc);
  }
  if (kWide)
    fe10_fold(r, acc);
  else
    fe10_fold32(r, acc);
  return r;
}

// Operands with F G < 6.4 (fe10_fold32); the looser always goes first.
__device__ __forceinline__ Fe10 fe10_mul(const Fe10 &f, const Fe10 &g) {
  return fe10_mul_t<false>(f, g);
}

// Any operands within fe10_mul's bounds (f up to 5 units, g up to 3.3).
__device__ __forceinline__ Fe10 fe10_mul_wide(const Fe10 &f, const Fe10 &g) {
  return fe10_mul_t<true>(f, g);
}
#else
// The host product: every column summed on its own, then one carry pass.  The
// same partial products and the same carries as the device form, so the two
// agree bit for bit.
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
      // an odd limb times an odd limb lands half a bit high: doubled
      if (j >= 0) {
        acc = mad64(((i & 1) && (j & 1)) ? f2[i] : f.v[i], g.v[j], acc);
      } else {
        const int jj = j + 10;
        acc = mad64(((i & 1) && (jj & 1)) ? f2[i] : f.v[i], g19[jj], acc);
      }
    }
    h[k] = acc;
  }
  return fe10_carry64(h);
}

OG_HD Fe10 fe10_mul_wide(const Fe10 &f, const Fe10 &g) { return fe10_mul(f, g); }
#endif

// The low bits of the canonical form of f*g from some of the product's ten
// columns, exact unless `exact` comes back false.  fe10_mul's result is
// V mod 2^255 + 19c for the column sums' integer V = sum S_k 2^a_k and
// c = floor(V / 2^255); columns 2..9 start at bit 51 or above, so its low
// word is S_0 + S_1 2^26 + 19c mod 2^32.  c is taken from columns 8 and 9
// alone, as floor(t / 2^51) for t = S_9 2^26 + S_8: columns 0..7 add below
// 2^244, under 2^40 in units of S_8's weight 2^204, so c can only be one more
// when t mod 2^51 is within 2^40 of 2^51.  The canonical form differs from
// the result (by -p, +19 in the low word) only when the result is at least
// p, which needs V mod 2^255 within 2^37 of 2^255, the same condition.  So
// the word is the canonical low word whenever t mod 2^51 is below
// 2^51 - 2^41, that is unless bits 15..24 of S_9 + floor(S_8 / 2^26) are all
// ones (about one product in 1,024); otherwise the canonical low word is the
// returned one plus 0, 19 or 38.  Same operand bounds as fe10_mul.
//
// Column 1 reaches only bits 26 and up, so fe10_mul_low26 leaves it out:
// S_0 + 19c, whose bits 0..25 are the canonical form's (the first three key
// bytes and the sign of x), from 30 of the 100 products; fe10_low_col1 adds
// the column for bits 26..31.  The key forms probe these bits and the search
// loop takes the sign of x from bit 0, forming the whole product only where
// `exact` is false or a probe passes.
template <int K>
OG_HD uint64_t fe10_low_column(const Fe10 &f, const Fe10 &g) {
  uint32_t a[10], b[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const int j = K - i;
    const int jj = j < 0 ? j + 10 : j;
    a[i] = ((i & 1) && (jj & 1)) ? 2u * f.v[i] : f.v[i];
    b[i] = j < 0 ? 19u * g.v[jj] : g.v[j];
  }
#ifdef __HIP_DEVICE_COMPILE__
  // ten v_mad_u64_u32 from zero: of columns 0 and 1 only the low word is
  // used, which the compiler would otherwise form with 32-bit multiplies and
  // adds, nineteen instructions
  return fe10_column0(a, b);
#else
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 10; ++i) acc = mad64(a[i], b[i], acc);
  return acc;
#endif
}

OG_HD uint32_t fe10_mul_low26(const Fe10 &f, const Fe10 &g, bool &exact) {
  const uint64_t acc9 = fe10_low_column<9>(f, g) + (fe10_low_column<8>(f, g) >> 26);
  exact = ((uint32_t)acc9 & (0x3ffu << 15)) != (0x3ffu << 15);
  return (uint32_t)fe10_low_column<0>(f, g) + 19u * (uint32_t)(acc9 >> 25);
}

// Column 1's part of the low word: fe10_mul_low26 + this is fe10_mul_low32.
OG_HD uint32_t fe10_low_col1(const Fe10 &f, const Fe10 &g) {
  return (uint32_t)fe10_low_column<1>(f, g) << 26;
}

OG_HD uint32_t fe10_mul_low32(const Fe10 &f, const Fe10 &g, bool &exact) {
  return fe10_mul_low26(f, g, exact) + fe10_low_col1(f, g);
}

constexpr uint32_t kLow26 = (1u << 26) - 1;

OG_HD Fe10 fe10_sq(const Fe10 &f) { return fe10_mul(f, f); }

OG_HD Fe10 fe10_add(const Fe10 &a, const Fe10 &b) {
  Fe10 r;
#pragma unroll
  for (int i = 0; i < 10; ++i) r.v[i] = a.v[i] + b.v[i];
  return r;
}

// a - b + 2p, limb by limb: every limb of 2p (2^27 - 38, 2^26 - 2, 2^27 - 2, ...)
// exceeds the corresponding limb of a tight b, so nothing goes negative.
OG_HD Fe10 fe10_sub(const Fe10 &a, const Fe10 &b) {
  Fe10 r;
  r.v[0] = a.v[0] + ((1u << 27) - 38u) - b.v[0];
#pragma unroll
  for (int i = 1; i < 10; ++i)
    r.v[i] = a.v[i] + ((i & 1) ? ((1u << 26) - 2u) : ((1u << 27) - 2u)) - b.v[i];
  return r;
}

// Tighten a loose value (limbs below 2^31).
OG_HD Fe10 fe10_carry(const Fe10 &a) {
  uint64_t h[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) h[i] = a.v[i];
  return fe10_carry64(h);
}

// The low 32 bits of the integer a tight value stands for.  Its canonical
// form is that integer, or that integer minus p (when it is at least p), whose
// low 32 bits are those of the integer plus 19: a filter on the leading 32
// bits of the key asks both and confirms on the canonical form.
OG_HD uint32_t fe10_low32(const Fe10 &a) { return a.v[0] + (a.v[1] << 26); }

// The canonical representative below p as eight 32-bit words.  q = [x >= p]
// comes from propagating x + 19 through the limbs; then x + 19q - 2^255 q is
// carried out exactly and packed (the technique of ref10's fe_tobytes).
//
// fe10_canon_product takes a value whose limbs are below 2^31 and whose
// integer is below 2p, with no carry pass first: the output of fe10_mul (below
// 2^255 + 2^44; only limb 1 can exceed its width, by at most 2^18).  The ripple
// for q then carries exactly whatever the limbs hold, q = floor((x + 19) /
// 2^255) is still [x >= p], and the final ripple normalises every limb.
// fe10_canon tightens an arbitrary loose value first.
OG_HD Fe fe10_canon_product(const Fe10 &t) {
  uint32_t q = (t.v[0] + 19u) >> 26;
#pragma unroll
  for (int i = 1; i < 10; ++i) q = (t.v[i] + q) >> fe10_width(i);
  uint32_t h[10];
  h[0] = t.v[0] + 19u * q;
#pragma unroll
  for (int i = 1; i < 10; ++i) h[i] = t.v[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int w = fe10_width(i);
    h[i + 1] += h[i] >> w;
    h[i] &= (1u << w) - 1;
  }
  h[9] &= (1u << 25) - 1;  // drops 2^255 q
  Fe r;
  r.v[0] = h[0] | (h[1] << 26);
  r.v[1] = (h[1] >> 6) | (h[2] << 19);
  r.v[2] = (h[2] >> 13) | (h[3] << 13);
  r.v[3] = (h[3] >> 19) | (h[4] << 6);
  r.v[4] = h[5] | (h[6] << 25);
  r.v[5] = (h[6] >> 7) | (h[7] << 19);
  r.v[6] = (h[7] >> 13) | (h[8] << 12);
  r.v[7] = (h[8] >> 20) | (h[9] << 6);
  return r;
}

OG_HD Fe fe10_canon(const Fe10 &a) { return fe10_canon_product(fe10_carry(a)); }

// Bit 0 of the canonical form of a product's output (the sign of x in a
// packed key), without forming it: the canonical value is x - qp with p odd,
// so its parity is that of x (bit 0 of limb 0, which no carry reaches) xor q.
OG_HD uint32_t fe10_parity_product(const Fe10 &t) {
  uint32_t q = (t.v[0] + 19u) >> 26;
#pragma unroll
  for (int i = 1; i < 10; ++i) q = (t.v[i] + q) >> fe10_width(i);
  return (t.v[0] ^ q) & 1u;
}

// Bit 0 of the canonical form of f*g (the sign of x in a packed key): from
// three columns where that settles it, else from the whole product.
OG_HD uint32_t fe10_parity_mul(const Fe10 &f, const Fe10 &g) {
  bool exact;
  const uint32_t lo = fe10_mul_low26(f, g, exact);
  uint32_t sign = lo & 1u;
  if (!exact) sign = fe10_parity_product(fe10_mul(f, g));
  return sign;
}

// The limbs of x mod 2^255 as they stand, with no reduction mod p: a value in
// [p, 2^255) keeps its non-canonical form (the field tests feed these to
// fe10_canon_product).
OG_HD Fe10 fe10_split(const Fe &c) {
  Fe10 r;
  r.v[0] = c.v[0] & 0x3ffffff;
  r.v[1] = ((c.v[0] >> 26) | (c.v[1] << 6)) & 0x1ffffff;
  r.v[2] = ((c.v[1] >> 19) | (c.v[2] << 13)) & 0x3ffffff;
  r.v[3] = ((c.v[2] >> 13) | (c.v[3] << 19)) & 0x1ffffff;
  r.v[4] = (c.v[3] >> 6) & 0x3ffffff;
  r.v[5] = c.v[4] & 0x1ffffff;
  r.v[6] = ((c.v[4] >> 25) | (c.v[5] << 7)) & 0x3ffffff;
  r.v[7] = ((c.v[5] >> 19) | (c.v[6] << 13)) & 0x1ffffff;
  r.v[8] = ((c.v[6] >> 12) | (c.v[7] << 20)) & 0x3ffffff;
  r.v[9] = (c.v[7] >> 6) & 0x1ffffff;
  return r;
}

// From any 8-word value below 2^256 (canonicalised first).
OG_HD Fe10 fe10_from(const Fe &x) { return fe10_split(fe_canon(x)); }

OG_HD Fe fe10_to(const Fe10 &a) { return fe10_canon(a); }

OG_HD Fe10 fe10_sqn(Fe10 a, int n) {
  for (int i = 0; i < n; ++i) a = fe10_sq(a);
  return a;
}

// x^(p-2): the ref10 chain, 254 squarings and 11 multiplications
// (src/field.rs:163-220).
OG_HD Fe10 fe10_invert(const Fe10 &z1) {
  Fe10 z2 = fe10_sq(z1);
  Fe10 z8 = fe10_sqn(z2, 2);
  Fe10 z9 = fe10_mul(z1, z8);
  Fe10 z11 = fe10_mul(z2, z9);
  Fe10 z22 = fe10_sq(z11);
  Fe10 z5 = fe10_mul(z9, z22);
  Fe10 z10 = fe10_mul(fe10_sqn(z5, 5), z5);
  Fe10 z20 = fe10_mul(fe10_sqn(z10, 10), z10);
  Fe10 z40 = fe10_mul(fe10_sqn(z20, 20), z20);
  Fe10 z50 = fe10_mul(fe10_sqn(z40, 10), z10);
  Fe10 z100 = fe10_mul(fe10_sqn(z50, 50), z50);
  Fe10 z200 = fe10_mul(fe10_sqn(z100, 100), z100);
  Fe10 t = fe10_mul(fe10_sqn(z200, 50), z50);
  return fe10_mul(fe10_sqn(t, 5), z11);
}

}  // namespace og
