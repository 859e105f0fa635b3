// The search kernels: one chain per lane, two candidates per table offset.
//
// Chain c of a search owns window w of 2H+1 consecutive candidates per round:
//     k = w*(2H+1) + H + j,  j = -H .. H,  centre P_w = P0 + (w(2H+1) + H) * 8G.
// For an affine offset Q_m = m*8G (m = 1..H) with dq = d*xq*yq, the unified
// Edwards addition gives both neighbours of a projective P = (X:Y:Z:T):
//     y(P + Q) = (Y*yq + X*xq) / (Z - T*dq),   x(P + Q) = (X*yq + Y*xq) / (Z + T*dq)
//     y(P - Q) = (Y*yq - X*xq) / (Z + T*dq),   x(P - Q) = (X*yq - Y*xq) / (Z - T*dq)
// (affine y3 = (y1y2 + x1x2)/(1 - d x1x2y1y2), multiplied through by Z).  The
// denominators never vanish on the curve (the formula is complete) and P is
// never normalised.  This is the reference's pair idea (src/pairs.rs:1-19,
// cuda-kernel/src/lib.rs:118-121) with the complete formula in place of its
// dedicated one.
//
// The pair shares one denominator, their product
//     E = (Z - T*dq)(Z + T*dq) = Z^2 + T^2 * (-dq^2),
// and each numerator is taken over E by the other denominator.  Multiplied
// out, both numerators are sums of the window's four products YZ, TX, XZ, TY
// (fixed for the window) times table constants:
//     y(P +- Q) = (alpha +- beta) / E,  alpha = YZ*yq + TX*(xq dq),  beta = XZ*xq + TY*(yq dq)
//     x(P +- Q) = (gamma +- delta) / E, gamma = XZ*yq - TY*(xq dq),  delta = YZ*xq - TX*(yq dq)
// so the table holds, per offset, xq, yq, dq, xq*dq, yq*dq, -dq^2, and the
// negations of xq*dq and yq*dq that make gamma and delta double products.
//
// Montgomery's trick runs over the E of each pair.  The forward pass forms
// E = ZZ + TT*(-dq^2) and the running product, and stores two field elements
// per pair: the prefix D before the pair and E.  After one inversion of the
// whole product (times Z, for the centre) the backward pass peels the
// inverses off (1/E = inv*D, inv *= E) and forms the two y from alpha and beta,
// each a double product by constants with one carry pass.  Per pair: one
// product by a constant and one multiplication forward; two multiplications,
// two double products by constants and two multiplications backward: 1,000
// partial products for two candidates, against 1,100 for the three
// constant products (T*dq, Y*yq, X*xq) and eight multiplications of the
// separate-denominator form.  A candidate costs 40 bytes of scratch written
// once and read once.  Keeping E costs one element per pair and saves
// recomputing it (the reference recomputes its denominators,
// src/pairs.rs:182-207, which suits a CPU's cache; here HBM has the room and
// the bandwidth, and the multiply issue rate is the bound).  The x coordinate
// of a candidate costs gamma and delta and one more multiplication, only where
// a filter needs the sign of x.
#pragma once
#include "fe10.hpp"
#include "fe_sginv.hpp"
#include "ge.hpp"
#include "keccak.hpp"
#include "og_scan.hpp"

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
  kModePattern = 4,  // general forms and expressions: the bank automaton of og_scan.hpp
};

// Scratch planes (field elements of 10 limbs): the chain's two per pair (D
// and E, planes 2m and 2m + 1), and for the pattern kernels the parking
// planes, which count down from plane 3H (kPlanesPerPair * H) to at least
// plane H, so that parked entries never reach a plane the backward pass has
// yet to read and planes 0 .. H - 1 stay free for the address kernel's
// derivation of deferred signs.  The key forms never park: 2H planes.
constexpr uint32_t kChainPlanes = 2;
constexpr uint32_t kPlanesPerPair = 3;
inline size_t scratch_planes(uint32_t half, bool pattern) {
  return pattern ? (size_t)kPlanesPerPair * half + 1 : (size_t)kChainPlanes * half;
}

// The offset table: per offset m (Q = (m + 1) * 8G), eight field elements of
// ten limbs.
constexpr uint32_t kTableTerms = 8;
constexpr uint32_t kTableWords = 10 * kTableTerms;
constexpr uint32_t kTabX = 0, kTabY = 10, kTabDq = 20, kTabXdq = 30, kTabYdq = 40, kTabNdq2 = 50,
                   kTabNxdq = 60, kTabNydq = 70;

// The derived constants of an offset: c[0..2] = xq, yq, dq (canonical) in,
// c[3..7] = xq*dq, yq*dq, -dq^2, -xq*dq, -yq*dq (canonical) out.  The negated
// products let gamma and delta (the x numerators) each be one double product.
OG_HD void offset_terms(Fe c[kTableTerms]) {
  c[3] = fe_canon(fe_mul(c[0], c[2]));
  c[4] = fe_canon(fe_mul(c[1], c[2]));
  c[5] = fe_canon(fe_sub(fe_zero(), fe_mul(c[2], c[2])));
  c[6] = fe_canon(fe_sub(fe_zero(), c[3]));
  c[7] = fe_canon(fe_sub(fe_zero(), c[4]));
}

constexpr int kMaxBuckets = 16;
constexpr int kMaxPatterns = 64;  // symbol-wise records; the banks bound the set
constexpr int kPatWords = 4 + 56;  // DevPattern: len, first, last, key_last, sets[56]
constexpr int kMaxRegexes = 16;
// The first automaton is kept in each workgroup's LDS when it has at most
// this many states (32 KiB).  A larger one is walked in global memory: a wave
// waits whenever any lane leaves the LDS part, so a partial copy does not pay.
constexpr uint32_t kLdsStates = 512;
// The pattern kernels' LDS for a filter: the first automaton (kLdsStates rows
// of 32 16-bit entries) or the window prefilter's two 2^18-bit maps.
constexpr uint32_t kPatternLdsWords = 2u << (18 - 5);
// The bitmap probe's LDS copy: at most 2^18 bits (32 KiB), so the key forms'
// three workgroups per CU hold three copies.
constexpr uint32_t kCoarseBits = 18;
constexpr uint32_t kCoarseWords = 1u << (kCoarseBits - 5);
// The key stage is skipped when at least this fraction of kNeedSamples
// pseudo-random keys would go to the address stage anyway.  On MI355X a key
// stage's automaton walk costs about 64 ps a candidate and the address stage
// about 107 ps (1,000 five-symbol words, 2^31-candidate launches), so skipping
// pays from about 40%; the bound leaves room for the sign of x, which every
// parked lane then forms.
constexpr uint32_t kNeedSamples = 4096;
constexpr double kParkAllRate = 0.9;
// The window prefilter in the key stage sends a key whose window hits both
// maps to the address stage; it is used while the expected share of such keys
// stays at most this (about 27 ps a candidate against 64 ps for one
// automaton's whole walk).
constexpr double kWindowParkMax = 0.25;

struct DevFilterArgs {
  uint32_t mode;
  uint32_t want[8], mask[8];      // kModePrefix
  uint32_t bits;                  // kModeBitmap
  const uint32_t *bitmap;
  uint32_t nbuckets;              // exact table behind the bitmap
  const uint32_t *buckets;        // [nbuckets][kBucketWords]
  const uint64_t *keys;
  const uint32_t *dirs;           //   the buckets' directories
  uint32_t dir_steps;             //   halvings after a directory lookup
  uint32_t nbanks;                // kModePattern: the automaton
  const uint32_t *banks;          // nbanks records of kBankWords
  uint32_t npat;                  // the same patterns, for the symbol-wise check
  const uint32_t *pats;           // npat records of kPatWords
  uint32_t nregex;                // regular expressions: their automata
  uint32_t regex_every;           //   (kModePattern) nonzero: one has no proxy in
                                  //   the banks, so every candidate runs them
  uint32_t key_proxy;             //   (key forms) the key tables hold the
                                  //   expressions' literal prefixes: a key hit
                                  //   only sends the lane to the automata
  const uint32_t *regex;          //   records, entries and tail-liveness
  const uint16_t *dfa;            //   bits: og_scan.hpp
  const uint32_t *dfa_end;        //   (the end flags of the automata's entries)
  const uint32_t *regex_live;
  uint32_t lds_states;            //   the first automaton's states, when it is
                                  //   copied to LDS whole (0: it is not)
  uint32_t defer_sign;            // (kModePattern, one-round launches) the key
                                  //   stage parks candidates without the sign
                                  //   of x, and the address kernel derives it
  uint32_t park_all;              // (kModePattern) the key stage would send
                                  //   nearly every candidate to the address
                                  //   stage: it is skipped, and every candidate
                                  //   the key forms leave goes there
  const uint32_t *hash;           // (kModePattern) a dictionary of plain
                                  //   substrings: the window prefilter's two
                                  //   maps (null: none)
  uint32_t hash_lmax;             //   its longest word
  uint32_t hash_ws;               //   its form, 8 x window length + stride
  uint32_t hash_key;              //   the key stage probes it
  uint32_t hash_addr;             //   the address stage probes it
  const uint32_t *coarse;         // (kModeBitmap) the bitmap or-folded to
  uint32_t coarse_bits;           //   coarse_bits <= kCoarseBits bits: each
                                  //   workgroup's LDS copy is probed first, and
                                  //   equals the map when coarse_bits == bits
};

struct SearchArgs {
  uint32_t *state;        // [4 coords][10 limbs][threads]
  const uint32_t *table;  // [half][kTableWords]
  uint32_t *scratch;      // [scratch_planes(half, pattern)][10][threads]
  uint32_t threads, half, rounds;
  uint64_t k_base;        // first candidate of this launch
  uint32_t stride[30];    // (y+x, y-x, 2dxy) of threads*(2H+1)*8G, Z = 1
  DevFilterArgs f;
  uint32_t *counter;      // survivors found (may exceed cap)
  HitRec *hits;
  uint32_t cap;
  uint32_t *parked;       // per chain: candidates left for the address kernel
                          // (one round per launch), or null: drained in place
};

struct SetupArgs {
  uint32_t *state;
  uint32_t threads;       // SoA stride
  uint32_t first;         // first chain to set up
  uint32_t count;         // chains to set up
  uint32_t half;
  uint64_t k0;            // candidate index of chain 0's first window
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
  const Fe10 a = fe10_mulc(fe10_add(p.Y, p.X), q.ypx);
  const Fe10 b = fe10_mulc(fe10_sub(p.Y, p.X), q.ymx);
  const Fe10 c = fe10_mulc(p.T, q.t2d);
  const Fe10 d = fe10_add(p.Z, p.Z);
  const Fe10 e = fe10_sub(a, b), h = fe10_add(a, b);
  const Fe10 f = fe10_sub(d, c), g = fe10_add(d, c);
  Point10 r;
  // f (up to 4 units) is the looser operand, so it goes first; operands up
  // to 4 x 3 units take the wide fold
  r.X = fe10_mul_wide(f, e);
  r.Y = fe10_mul_wide(g, h);
  r.Z = fe10_mul_wide(f, g);
  r.T = fe10_mul_wide(e, h);
  return r;
}

}  // namespace og
