// ed25519 group arithmetic for the search: extended coordinates
// (x = X/Z, y = Y/Z, T = XY/Z), the formulas of the reference's src/curve.rs:130-198.
//
// Three shapes are used:
//   * Point      the running base point of a chain, extended projective;
//   * Cached     an addend in the form (y+x, y-x, Z, 2d*T) of src/curve.rs:177-184;
//   * Cached1    the same with Z scaled to one on the host, so an addition
//                costs seven multiplications instead of eight;
//   * TableQ     an affine offset Q with (x, y, d*x*y): the y coordinate of
//                P + Q and P - Q for a projective P then needs three
//                multiplications for both (see pair_terms below).
#pragma once
#include "fe.hpp"

namespace og {

struct Point {
  Fe X, Y, Z, T;
};

struct Cached {
  Fe ypx, ymx, Z, t2d;
};

struct Cached1 {
  Fe ypx, ymx, t2d;
};

OG_HD Point pt_identity() {
  Point p;
  p.X = fe_zero();
  p.Y = fe_one();
  p.Z = fe_one();
  p.T = fe_zero();
  return p;
}

// P + Q for a cached Q with Z = 1 (src/curve.rs:130-145 add, :168-175 to_p3;
// the Z2 = 1 shortcut is cuda-kernel/src/curve.rs:42-67's).
OG_HD Point pt_add1(const Point &p, const Cached1 &q) {
  Fe a = fe_mul(fe_add(p.Y, p.X), q.ypx);
  Fe b = fe_mul(fe_sub(p.Y, p.X), q.ymx);
  Fe c = fe_mul(p.T, q.t2d);
  Fe d = fe_add(p.Z, p.Z);
  Fe e = fe_sub(a, b), h = fe_add(a, b);
  Fe f = fe_sub(d, c), g = fe_add(d, c);
  Point r;
  r.X = fe_mul(e, f);
  r.Y = fe_mul(g, h);
  r.Z = fe_mul(f, g);
  r.T = fe_mul(e, h);
  return r;
}

// P + Q for a general cached Q.
OG_HD Point pt_add(const Point &p, const Cached &q) {
  Fe a = fe_mul(fe_add(p.Y, p.X), q.ypx);
  Fe b = fe_mul(fe_sub(p.Y, p.X), q.ymx);
  Fe c = fe_mul(p.T, q.t2d);
  Fe zz = fe_mul(p.Z, q.Z);
  Fe d = fe_add(zz, zz);
  Fe e = fe_sub(a, b), h = fe_add(a, b);
  Fe f = fe_sub(d, c), g = fe_add(d, c);
  Point r;
  r.X = fe_mul(e, f);
  r.Y = fe_mul(g, h);
  r.Z = fe_mul(f, g);
  r.T = fe_mul(e, h);
  return r;
}

// 2P (src/curve.rs:148-164, then to_p3).
OG_HD Point pt_dbl(const Point &p) {
  Fe xx = fe_sq(p.X);
  Fe yy = fe_sq(p.Y);
  Fe zz2 = fe_sq(p.Z);
  zz2 = fe_add(zz2, zz2);
  Fe s = fe_sq(fe_add(p.X, p.Y));
  Fe y = fe_add(yy, xx);
  Fe z = fe_sub(yy, xx);
  Fe x = fe_sub(s, y);
  Fe t = fe_sub(zz2, z);
  Point r;
  r.X = fe_mul(x, t);
  r.Y = fe_mul(y, z);
  r.Z = fe_mul(z, t);
  r.T = fe_mul(x, y);
  return r;
}

OG_HD Cached pt_to_cached(const Point &p, const Fe &two_d) {
  Cached c;
  c.ypx = fe_add(p.Y, p.X);
  c.ymx = fe_sub(p.Y, p.X);
  c.Z = p.Z;
  c.t2d = fe_mul(p.T, two_d);
  return c;
}

// m * Q for a cached Q, m below 2^64, by double-and-add from the top bit.
OG_HD Point pt_mul_u64(uint64_t m, const Cached &q) {
  Point acc = pt_identity();
  for (int i = 63; i >= 0; --i) {
    acc = pt_dbl(acc);
    if ((m >> i) & 1) acc = pt_add(acc, q);
  }
  return acc;
}

}  // namespace og
