// The search kernel: one chain per lane, two candidates per table offset.
//
// Chain g owns window w of 2H+1 consecutive candidates per round:
//     k = w*(2H+1) + H + j,   j = -H .. H,   window centre P_w = P0 + (w(2H+1)+H) * 8G.
// For an affine offset Q_m = m*8G (m = 1..H) with precomputed dq = d*xq*yq, the
// unified Edwards addition gives the y coordinates of both neighbours of a
// projective P = (X:Y:Z:T):
//     y(P + Q) = (Y*yq + X*xq) / (Z - T*dq)
//     y(P - Q) = (Y*yq - X*xq) / (Z + T*dq)
// (affine y3 = (y1y2 + x1x2)/(1 - d x1x2y1y2), multiplied through by Z).  Three
// multiplications serve two candidates, the denominators never vanish on the
// curve (the formula is complete), and P never has to be normalised.  This is
// the reference's pair idea (src/pairs.rs:1-19, cuda-kernel/src/lib.rs:118-121)
// with the complete formula in place of its dedicated one.
//
// Montgomery's trick is folded (src/pairs.rs:182-207): the forward pass stores
// s_i = n_i * (d_0...d_{i-1}) and keeps only the running product; after one
// inversion the backward pass recomputes each d_i from P and the table (one
// multiplication) instead of storing it, so a candidate costs 32 bytes of
// scratch written once and read once.  Per pair: 7 mul forward, 5 backward.
#pragma once
#include "ge.hpp"
#include "keccak.hpp"

namespace og {

// Hit record: candidate index and canonical y (sign clear).
struct HitRec {
  uint64_t k;
  uint32_t y[8];
};

enum : uint32_t {
  kModePrefix = 1,   // masked compare of the key words (one literal prefix)
  kModeBitmap = 2,   // prefix bitmap, then the exact table when present
  kModePattern = 4,  // general forms on the full address (sign + SHA3-256)
};

constexpr int kMaxBuckets = 16;
constexpr int kMaxPatterns = 8;

struct DevFilterArgs {
  uint32_t want[8], mask[8];      // kModePrefix
  uint32_t bits;                  // kModeBitmap
  const uint32_t *bitmap;
  uint32_t nbuckets;              // exact table behind the bitmap
  uint32_t boff[kMaxBuckets], bcnt[kMaxBuckets];
  uint64_t bmask[kMaxBuckets];
  const uint64_t *keys;
  uint32_t npat;                  // kModePattern
  const uint32_t *pats;           // DevPattern records, 4 + 56 words each
};

struct SearchArgs {
  uint32_t *state;        // [4 coords][8 limbs][threads]
  const uint32_t *table;  // [half][xq, yq, dq][8]
  uint32_t *scratch;      // [2*half][8][threads]
  uint32_t threads, half, rounds;
  uint64_t k_base;        // first candidate of this launch
  uint32_t stride[24];    // (y+x, y-x, 2dxy) of threads*(2H+1)*8G, Z = 1
  DevFilterArgs f;
  uint32_t *counter;      // survivors found (may exceed cap)
  HitRec *hits;
  uint32_t cap;
};

struct SetupArgs {
  uint32_t *state;
  uint32_t threads, half;
  uint32_t start[32];     // P0 as X, Y, Z, T
  uint32_t eight[32];     // 8G cached: y+x, y-x, Z, 2dT
  uint32_t two_d[8];
};

}  // namespace og
