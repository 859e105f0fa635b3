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

// The two checksum bytes of a packed key given as eight little-endian words.
// Returns checksum[0] | checksum[1] << 8.
OG_HD uint32_t onion_checksum(const uint32_t pk[8]) {
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
