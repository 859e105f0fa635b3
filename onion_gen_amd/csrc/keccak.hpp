// Keccak-f[1600] and the one SHA3-256 the address needs (FIPS 202), for the
// host and for gfx950.  The address checksum is
//   SHA3-256(".onion checksum" || pk || 0x03)[0..2]            (src/key.rs:106-113)
// a 48-byte message: one block of the 136-byte rate, so one permutation.
#pragma once
#include "fe.hpp"

namespace og {

OG_HD uint64_t rotl64(uint64_t x, int n) { return n == 0 ? x : (x << n) | (x >> (64 - n)); }

#define OG_KECCAK_RC                                                                    \
  {0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull,                 \
   0x8000000080008000ull, 0x000000000000808bull, 0x0000000080000001ull,                 \
   0x8000000080008081ull, 0x8000000000008009ull, 0x000000000000008aull,                 \
   0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,                 \
   0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull,                 \
   0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,                 \
   0x000000000000800aull, 0x800000008000000aull, 0x8000000080008081ull,                 \
   0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull}

// The round constants: in constant memory for the device (the round index is
// wave-uniform, so each is one scalar load), a static table for the host.  A
// local array indexed by the loop counter would live in scratch memory.
static __constant__ uint64_t kKeccakRcDev[24] = OG_KECCAK_RC;
static const uint64_t kKeccakRcHost[24] = OG_KECCAK_RC;

OG_HD uint64_t keccak_rc(int round) {
#ifdef __HIP_DEVICE_COMPILE__
  return kKeccakRcDev[round];
#else
  return kKeccakRcHost[round];
#endif
}

// The 24 rounds over a 5x5 lane state a[x + 5y].  Rho offsets and the pi
// permutation are written out so that every index is a compile-time constant
// and the state stays in registers.
OG_HD void keccak_f1600(uint64_t a[25]) {
#pragma unroll 1
  for (int round = 0; round < 24; ++round) {
    uint64_t c[5], d[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
    for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
#pragma unroll
    for (int i = 0; i < 25; ++i) a[i] ^= d[i % 5];
    // rho + pi: b[y + 5*((2x + 3y) % 5)] = rot(a[x + 5y], r[x][y])
    b[0] = a[0];
    b[10] = rotl64(a[1], 1);
    b[20] = rotl64(a[2], 62);
    b[5] = rotl64(a[3], 28);
    b[15] = rotl64(a[4], 27);
    b[16] = rotl64(a[5], 36);
    b[1] = rotl64(a[6], 44);
    b[11] = rotl64(a[7], 6);
    b[21] = rotl64(a[8], 55);
    b[6] = rotl64(a[9], 20);
    b[7] = rotl64(a[10], 3);
    b[17] = rotl64(a[11], 10);
    b[2] = rotl64(a[12], 43);
    b[12] = rotl64(a[13], 25);
    b[22] = rotl64(a[14], 39);
    b[23] = rotl64(a[15], 41);
    b[8] = rotl64(a[16], 45);
    b[18] = rotl64(a[17], 15);
    b[3] = rotl64(a[18], 21);
    b[13] = rotl64(a[19], 8);
    b[14] = rotl64(a[20], 18);
    b[24] = rotl64(a[21], 2);
    b[9] = rotl64(a[22], 61);
    b[19] = rotl64(a[23], 56);
    b[4] = rotl64(a[24], 14);
    // chi
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
#pragma unroll
      for (int x = 0; x < 5; ++x) a[y + x] = b[y + x] ^ (~b[y + (x + 1) % 5] & b[y + (x + 2) % 5]);
    }
    a[0] ^= keccak_rc(round);
  }
}

#ifdef __HIP_DEVICE_COMPILE__
// The checksum permutation for gfx950: each 64-bit lane as two 32-bit halves,
// rotations as two v_alignbit_b32 (funnel shifts), chi and the five-way column
// XOR as v_bitop3_b32 (any three-input bitwise function), and a last round cut
// to the one lane the checksum reads.  About 180 VALU instructions a round
// where the 64-bit form compiles to 288.
struct KLane {
  uint32_t lo, hi;
};

// rotl by a compile-time amount (1..63, not 32: no rho offset is 0 mod 32
// besides lane 0's, which is not rotated).
template <int N>
__device__ __forceinline__ KLane krotl(KLane x) {
  static_assert(N > 0 && N < 64 && N != 32, "rotation");
  if (N < 32)
    return {__builtin_amdgcn_alignbit(x.lo, x.hi, 32 - N), __builtin_amdgcn_alignbit(x.hi, x.lo, 32 - N)};
  return {__builtin_amdgcn_alignbit(x.hi, x.lo, 64 - N), __builtin_amdgcn_alignbit(x.lo, x.hi, 64 - N)};
}

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0x96);  // a ^ b ^ c
}

__device__ __forceinline__ uint32_t chi3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0xd2);  // a ^ (~b & c)
}

// Theta: the column parities c and their rotations r.  Lane a[x + 5y] then
// takes a ^ c[x-1] ^ rotl1(c[x+1]) as one three-input XOR a half, rather than
// forming that sum once per column and XORing it in (10 fewer instructions a
// round).
struct KTheta {
  KLane c[5], r[5];
};

__device__ __forceinline__ KTheta ktheta(const KLane a[25]) {
  KTheta t;
#pragma unroll
  for (int x = 0; x < 5; ++x) {
    t.c[x].lo = xor3(xor3(a[x].lo, a[x + 5].lo, a[x + 10].lo), a[x + 15].lo, a[x + 20].lo);
    t.c[x].hi = xor3(xor3(a[x].hi, a[x + 5].hi, a[x + 10].hi), a[x + 15].hi, a[x + 20].hi);
  }
#pragma unroll
  for (int x = 0; x < 5; ++x) t.r[x] = krotl<1>(t.c[x]);
  return t;
}

// Lane a of column x after theta.
__device__ __forceinline__ KLane ktd(const KLane &a, const KTheta &t, int x) {
  const KLane &c = t.c[(x + 4) % 5], &r = t.r[(x + 1) % 5];
  return {xor3(a.lo, c.lo, r.lo), xor3(a.hi, c.hi, r.hi)};
}

// One round of Keccak-f[1600] on the split-lane state.
__device__ __forceinline__ void kround(KLane a[25], uint64_t rc) {
  KLane b[25];
  const KTheta t = ktheta(a);
  b[0] = ktd(a[0], t, 0);
  b[10] = krotl<1>(ktd(a[1], t, 1));
  b[20] = krotl<62>(ktd(a[2], t, 2));
  b[5] = krotl<28>(ktd(a[3], t, 3));
  b[15] = krotl<27>(ktd(a[4], t, 4));
  b[16] = krotl<36>(ktd(a[5], t, 0));
  b[1] = krotl<44>(ktd(a[6], t, 1));
  b[11] = krotl<6>(ktd(a[7], t, 2));
  b[21] = krotl<55>(ktd(a[8], t, 3));
  b[6] = krotl<20>(ktd(a[9], t, 4));
  b[7] = krotl<3>(ktd(a[10], t, 0));
  b[17] = krotl<10>(ktd(a[11], t, 1));
  b[2] = krotl<43>(ktd(a[12], t, 2));
  b[12] = krotl<25>(ktd(a[13], t, 3));
  b[22] = krotl<39>(ktd(a[14], t, 4));
  b[23] = krotl<41>(ktd(a[15], t, 0));
  b[8] = krotl<45>(ktd(a[16], t, 1));
  b[18] = krotl<15>(ktd(a[17], t, 2));
  b[3] = krotl<21>(ktd(a[18], t, 3));
  b[13] = krotl<8>(ktd(a[19], t, 4));
  b[14] = krotl<18>(ktd(a[20], t, 0));
  b[24] = krotl<2>(ktd(a[21], t, 1));
  b[9] = krotl<61>(ktd(a[22], t, 2));
  b[19] = krotl<56>(ktd(a[23], t, 3));
  b[4] = krotl<14>(ktd(a[24], t, 4));
#pragma unroll
  for (int y = 0; y < 25; y += 5) {
#pragma unroll
    for (int x = 0; x < 5; ++x) {
      a[y + x].lo = chi3(b[y + x].lo, b[y + (x + 1) % 5].lo, b[y + (x + 2) % 5].lo);
      a[y + x].hi = chi3(b[y + x].hi, b[y + (x + 1) % 5].hi, b[y + (x + 2) % 5].hi);
    }
  }
  a[0].lo ^= (uint32_t)rc;
  a[0].hi ^= (uint32_t)(rc >> 32);
}

__device__ __forceinline__ uint32_t onion_checksum_gfx950(const uint32_t pk[8]) {
  KLane a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = {0u, 0u};
  // message bytes: ".onion checksum" (15) | pk (32) | 0x03, then SHA3 padding
  // 0x06 at byte 48 and 0x80 at byte 135; lane i holds bytes 8i..8i+7.
  a[0] = {0x696e6f2eu, 0x63206e6fu};                     // ".onion c"
  a[1] = {0x6b636568u, 0x006d7573u | (pk[0] << 24)};     // "hecksum" + pk[0]
#pragma unroll
  for (int i = 2; i < 5; ++i)
    a[i] = {__builtin_amdgcn_alignbit(pk[2 * i - 3], pk[2 * i - 4], 8),
            __builtin_amdgcn_alignbit(pk[2 * i - 2], pk[2 * i - 3], 8)};
  a[5] = {__builtin_amdgcn_alignbit(pk[7], pk[6], 8), (pk[7] >> 8) | (0x03u << 24)};
  a[6] = {0x06u, 0u};
  a[16] = {0u, 0x80000000u};
  // Round 0 on its own: 17 of its lanes are zero and three constant, which
  // the compiler folds.
  kround(a, keccak_rc(0));
  // The table's address taken once, in scalar registers: left to itself the
  // compiler reloads it from the global offset table every round, a scalar
  // load the round's constant then waits on.
  typedef const __attribute__((address_space(4))) uint64_t *cptr64;
  cptr64 rc = (cptr64)kKeccakRcDev;
  asm volatile("" : "+s"(rc));
#pragma unroll 1
  for (int round = 1; round < 23; ++round) kround(a, rc[round]);
  // round 23, cut to the low half of lane 0: chi of b[0], b[1], b[2] and iota
  const KTheta t = ktheta(a);
  const uint32_t b0 = ktd(a[0], t, 0).lo;
  const uint32_t b1 = krotl<44>(ktd(a[6], t, 1)).lo;
  const uint32_t b2 = krotl<43>(ktd(a[12], t, 2)).lo;
  const uint32_t lane0 = chi3(b0, b1, b2) ^ (uint32_t)keccak_rc(23);
  return lane0 & 0xffffu;
}
#endif

// The two checksum bytes of a packed key given as eight little-endian words.
// Returns checksum[0] | checksum[1] << 8.
OG_HD uint32_t onion_checksum(const uint32_t pk[8]) {
#ifdef __HIP_DEVICE_COMPILE__
  return onion_checksum_gfx950(pk);
#endif
  // message bytes: ".onion checksum" (15) | pk (32) | 0x03, then SHA3 padding
  // 0x06 at byte 48 and 0x80 at byte 135 (the last byte of the rate).
  uint64_t a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = 0;
  // ".onion c" and "hecksum" as little-endian lanes
  const uint64_t l0 = 0x63206e6f696e6f2eull;  // ".onion c"
  const uint64_t l1 = 0x006d75736b636568ull;  // "hecksum" + byte 15 from pk
  // bytes 15..46 carry pk[0..31]: lane i (bytes 8i..8i+7) takes pk bytes 8i-15 ..
  uint64_t w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) w[i] = (uint64_t)pk[2 * i] | ((uint64_t)pk[2 * i + 1] << 32);
  a[0] = l0;
  a[1] = l1 | (w[0] << 56);
  a[2] = (w[0] >> 8) | (w[1] << 56);
  a[3] = (w[1] >> 8) | (w[2] << 56);
  a[4] = (w[2] >> 8) | (w[3] << 56);
  a[5] = (w[3] >> 8) | (0x03ull << 56);
  a[6] = 0x06ull;
  a[16] = 0x8000000000000000ull;
  keccak_f1600(a);
  return (uint32_t)(a[0] & 0xffffu);
}

}  // namespace og
