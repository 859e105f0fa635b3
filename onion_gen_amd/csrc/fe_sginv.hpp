// Inversion in GF(2^255 - 19) by Bernstein and Yang's divstep recurrence
// ("Fast constant-time gcd computation and modular inversion", 2019), in its
// half-delta form, over signed 30-bit limbs with 32-bit words.  It replaces
// the ref10 exponentiation chain (src/field.rs:163-220: 254 squarings and 11
// multiplications, about 33,000 VALU instructions in the 10-limb product)
// where one inversion serves a search window: 20 batches of 30 divsteps, each
// batch run on the low words of f and g alone and then applied to the whole f,
// g (exact division by 2^30) and to the coefficients d, e (mod p, the
// division made exact by adding a multiple of p).
//
// Invariants (f = p, g = x, d = 0, e = 1 at the start): f = d x and g = e x
// (mod p); f odd; d and e in (-2p, p) with limbs 0..7 in [0, 2^30) and a signed
// limb 8.  600 divsteps bring g to 0 and f to +-1 for every input below 2^256
// (the recurrence's bound is 590), so x^-1 = +-d.  The same steps run for
// every input: the work is fixed, as a SIMT wave needs.
#pragma once
#include "fe10.hpp"

namespace og {

struct Sg9 {
  int32_t v[9];  // value sum v[i] 2^(30 i)
};

constexpr int32_t kSgMask = (1 << 30) - 1;
constexpr uint32_t kSgPinv30 = 395589093u;  // p^-1 mod 2^30
OG_HD constexpr int32_t sg_p30(int i) {      // the limbs of p = 2^255 - 19
  return i == 0 ? (1 << 30) - 19 : (i == 8 ? (1 << 15) - 1 : (1 << 30) - 1);
}

// 30 divsteps on the low words of f (odd) and g.  zeta = -(delta + 1/2)
// carries the step state; the returned matrix (u, v, q, r) maps (f, g) of
// the batch's start to 2^30 times (f, g) at its end.  |u| + |v| and |q| + |r|
// stay at most 2^30.
OG_HD int32_t sg_divsteps30(int32_t zeta, uint32_t f, uint32_t g, int32_t t[4]) {
  uint32_t u = 1, v = 0, q = 0, r = 1;
#pragma unroll
  for (int i = 0; i < 30; ++i) {
    const uint32_t c1 = (uint32_t)(zeta >> 31);  // all ones when delta > 0
    const uint32_t c2 = 0u - (g & 1u);           // all ones when g is odd
    // g odd: g += f, or g -= f when delta > 0 (and then f, g trade places)
    g += ((f ^ c1) - c1) & c2;
    q += ((u ^ c1) - c1) & c2;
    r += ((v ^ c1) - c1) & c2;
    const uint32_t c3 = c1 & c2;
    zeta = (int32_t)(((uint32_t)zeta ^ c3) - 1u);  // delta -> 1 - delta, else 1 + delta
    f += g & c3;
    u += q & c3;
    v += r & c3;
    g >>= 1;
    u <<= 1;
    v <<= 1;
  }
  t[0] = (int32_t)u;
  t[1] = (int32_t)v;
  t[2] = (int32_t)q;
  t[3] = (int32_t)r;
  return zeta;
}

// (f, g) <- (u f + v g, q f + r g) / 2^30, exact: the batch made the low 30
// bits of both sums zero.
OG_HD void sg_update_fg(Sg9 &f, Sg9 &g, const int32_t t[4]) {
  const int64_t u = t[0], v = t[1], q = t[2], r = t[3];
  int64_t cf = u * f.v[0] + v * g.v[0], cg = q * f.v[0] + r * g.v[0];
  cf >>= 30;
  cg >>= 30;
#pragma unroll
  for (int i = 1; i < 9; ++i) {
    cf += u * f.v[i] + v * g.v[i];
    cg += q * f.v[i] + r * g.v[i];
    f.v[i - 1] = (int32_t)cf & kSgMask;
    g.v[i - 1] = (int32_t)cg & kSgMask;
    cf >>= 30;
    cg >>= 30;
  }
  f.v[8] = (int32_t)cf;
  g.v[8] = (int32_t)cg;
}

// (d, e) <- (u d + v e, q d + r e) / 2^30 mod p: md and me multiples of p
// (first u or v for a negative d or e, which keeps the result in (-2p, p),
// then the amount that clears the low 30 bits) make the division exact.
OG_HD void sg_update_de(Sg9 &d, Sg9 &e, const int32_t t[4]) {
  const int32_t u = t[0], v = t[1], q = t[2], r = t[3];
  const int32_t sd = d.v[8] >> 31, se = e.v[8] >> 31;
  int32_t md = (u & sd) + (v & se), me = (q & sd) + (r & se);
  int64_t cd = (int64_t)u * d.v[0] + (int64_t)v * e.v[0];
  int64_t ce = (int64_t)q * d.v[0] + (int64_t)r * e.v[0];
  md -= (int32_t)((kSgPinv30 * (uint32_t)cd + (uint32_t)md) & (uint32_t)kSgMask);
  me -= (int32_t)((kSgPinv30 * (uint32_t)ce + (uint32_t)me) & (uint32_t)kSgMask);
  cd += (int64_t)sg_p30(0) * md;
  ce += (int64_t)sg_p30(0) * me;
  cd >>= 30;
  ce >>= 30;
#pragma unroll
  for (int i = 1; i < 9; ++i) {
    cd += (int64_t)u * d.v[i] + (int64_t)v * e.v[i] + (int64_t)sg_p30(i) * md;
    ce += (int64_t)q * d.v[i] + (int64_t)r * e.v[i] + (int64_t)sg_p30(i) * me;
    d.v[i - 1] = (int32_t)cd & kSgMask;
    e.v[i - 1] = (int32_t)ce & kSgMask;
    cd >>= 30;
    ce >>= 30;
  }
  d.v[8] = (int32_t)cd;
  e.v[8] = (int32_t)ce;
}

// Limbs 0..7 back to [0, 2^30), the sign in limb 8.
OG_HD void sg_carry(Sg9 &a) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int32_t c = a.v[i] >> 30;
    a.v[i] &= kSgMask;
    a.v[i + 1] += c;
  }
}

// a + p where a is negative (mask all ones), else a.
OG_HD void sg_add_p_if(Sg9 &a, int32_t mask) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a.v[i] += sg_p30(i) & mask;
  sg_carry(a);
}

OG_HD Sg9 sg_from(const Fe &x) {
  Sg9 a;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int bit = 30 * i, w = bit >> 5, sh = bit & 31;
    uint32_t lo = x.v[w] >> sh;
    if (sh > 2 && w + 1 < 8) lo |= x.v[w + 1] << (32 - sh);
    a.v[i] = (int32_t)(lo & (uint32_t)kSgMask);
  }
  return a;
}

// A value in [0, p) as eight 32-bit words.
OG_HD Fe sg_to(const Sg9 &a) {
  Fe r;
  uint64_t acc = 0;
  int bits = 0, j = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    acc |= (uint64_t)(uint32_t)a.v[i] << bits;
    bits += 30;
    if (bits >= 32) {
      r.v[j++] = (uint32_t)acc;
      acc >>= 32;
      bits -= 32;
    }
  }
  return r;
}

// x^-1 for a canonical x (0 maps to 0).
OG_HD Fe fe_invert_sg(const Fe &x) {
  Sg9 f, g = sg_from(x), d, e;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f.v[i] = sg_p30(i);
    d.v[i] = 0;
    e.v[i] = 0;
  }
  e.v[0] = 1;
  int32_t zeta = -1;  // delta = 1/2
#pragma unroll 1
  for (int it = 0; it < 20; ++it) {
    int32_t t[4];
    zeta = sg_divsteps30(zeta, (uint32_t)f.v[0], (uint32_t)g.v[0], t);
    sg_update_fg(f, g, t);
    sg_update_de(d, e, t);
  }
  // f = +-1 (or p for x = 0, with d = 0): x^-1 = sign(f) d, brought into [0, p)
  sg_add_p_if(d, d.v[8] >> 31);  // (-2p, p) -> (-p, p)
  const int32_t neg = f.v[8] >> 31;
#pragma unroll
  for (int i = 0; i < 9; ++i) d.v[i] = (d.v[i] ^ neg) - neg;
  sg_carry(d);
  sg_add_p_if(d, d.v[8] >> 31);  // -> [0, p)
  return sg_to(d);
}

// The same for a 10-limb value (any loose value within fe10_carry's bounds).
OG_HD Fe10 fe10_invert_sg(const Fe10 &z) { return fe10_split(fe_invert_sg(fe10_canon(z))); }

}  // namespace og
