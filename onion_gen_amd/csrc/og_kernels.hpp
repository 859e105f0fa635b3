// The search kernels: one chain per lane, two candidates per table offset.
//
// Chain c of a search owns window w of 2H+1 consecutive candidates per round:
//     k = w*(2H+1) + H + j,  j = -H .. H,  centre P_w = P0 + (w(2H+1) + H) * 8G.
// For an affine offset Q_m = m*8G (m = 1..H) with precomputed dq = d*xq*yq, the
// unified Edwards addition gives the y coordinates of both neighbours of a
// projective P = (X:Y:Z:T):
//     y(P + Q) = (Y*yq + X*xq) / (Z - T*dq)
//     y(P - Q) = (Y*yq - X*xq) / (Z + T*dq)
// (affine y3 = (y1y2 + x1x2)/(1 - d x1x2y1y2), multiplied through by Z).  Three
// multiplications serve two candidates, the denominators never vanish on the
// curve (the formula is complete) and P is never normalised.  This is the
// reference's pair idea (src/pairs.rs:1-19, cuda-kernel/src/lib.rs:118-121)
// with the complete formula in place of its dedicated one.
//
// Montgomery's trick is folded (src/pairs.rs:182-207): the forward pass stores
// s_i = n_i * (d_0...d_{i-1}) and keeps only the running product; after one
// inversion the backward pass recomputes each d_i from P and the table (one
// multiplication) instead of storing it, so a candidate costs one field element
// of scratch written once and read once.  Per pair: 7 mul forward, 5 backward.
// The x coordinates (needed only by the checksum forms) share the denominators:
// x(P + Q) = (X yq + Y xq)/(Z + T dq) has the denominator of y(P - Q).
#pragma once
#include "fe10.hpp"
#include "ge.hpp"
#include "keccak.hpp"

namespace og {

// Hit record: candidate index, search slot and canonical y (sign clear).
struct HitRec {
  uint64_t k;
  uint32_t slot;
  uint32_t y[8];
  uint32_t pad;
};

enum : uint32_t {
  kModePrefix = 1,   // masked compare of the key words (one literal prefix)
  kModeBitmap = 2,   // prefix bitmap, then the exact table when present
  kModePattern = 4,  // general forms on the full address (sign + SHA3-256)
};

constexpr int kMaxBuckets = 16;
constexpr int kMaxPatterns = 8;
constexpr int kPatWords = 4 + 56;  // DevPattern: len, first, last, key_last, sets[56]

struct DevFilterArgs {
  uint32_t mode;
  uint32_t want[8], mask[8];      // kModePrefix
  uint32_t bits;                  // kModeBitmap
  const uint32_t *bitmap;
  uint32_t nbuckets;              // exact table behind the bitmap
  const uint32_t *buckets;        // [nbuckets][offset, count, mask_lo, mask_hi]
  const uint64_t *keys;
  uint32_t npat;                  // kModePattern
  const uint32_t *pats;           // npat records of kPatWords
};

struct SearchArgs {
  uint32_t *state;        // [4 coords][10 limbs][threads]
  const uint32_t *table;  // [half][xq, yq, dq][10]
  uint32_t *scratch;      // [2*half][10][threads], then the same again for x
  uint32_t threads, half, rounds;
  uint64_t k_base;        // first candidate of this launch
  uint32_t stride[30];    // (y+x, y-x, 2dxy) of threads*(2H+1)*8G, Z = 1
  DevFilterArgs f;
  uint32_t *counter;      // survivors found (may exceed cap)
  HitRec *hits;
  uint32_t cap;
};

struct SetupArgs {
  uint32_t *state;
  uint32_t threads;       // SoA stride
  uint32_t first;         // first chain to set up
  uint32_t count;         // chains to set up
  uint32_t half;
  uint32_t start[32];     // P0 as X, Y, Z, T (8 x 32)
  uint32_t eight[32];     // 8G cached: y+x, y-x, Z, 2dT
  uint32_t two_d[8];
};

struct Point10 {
  Fe10 X, Y, Z, T;
};

struct Cached10 {
  Fe10 ypx, ymx, t2d;  // Z = 1
};

// P + Q for a cached Q with Z = 1, seven multiplications (src/curve.rs:130-175).
__device__ __forceinline__ Point10 pt10_add1(const Point10 &p, const Cached10 &q) {
  const Fe10 a = fe10_mul(fe10_add(p.Y, p.X), q.ypx);
  const Fe10 b = fe10_mul(fe10_sub(p.Y, p.X), q.ymx);
  const Fe10 c = fe10_mul(p.T, q.t2d);
  const Fe10 d = fe10_add(p.Z, p.Z);
  const Fe10 e = fe10_sub(a, b), h = fe10_add(a, b);
  const Fe10 f = fe10_sub(d, c), g = fe10_add(d, c);
  Point10 r;
  // f (up to 4 units) is the looser operand, so it goes first
  r.X = fe10_mul(f, e);
  r.Y = fe10_mul(g, h);
  r.Z = fe10_mul(f, g);
  r.T = fe10_mul(e, h);
  return r;
}

}  // namespace og
