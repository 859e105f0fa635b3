// The hit sink: survivors become verified key records, in batches.
//
// For every survivor (offset 8k, packed y) of a search seeded by `seed`:
//   1. the candidate's secret: the seed's clamped scalar plus the offset, the
//      right half of the expanded secret carried over, and the clamping check
//      (src/key.rs:167-206 expanded_secret_at_offset, add_to_scalar,
//      clamping_intact): a hit whose offset broke clamping is dropped;
//   2. its public key re-derived from that scalar alone, on the device, for the
//      whole batch at once (the independent check of src/run.rs:1089-1104 and
//      src/verify.rs:89-189, which the reference does per hit on the host);
//   3. the 255 bits of y compared with the survivor's, the sign taken from the
//      re-derived key, and the hostname built (src/key.rs:121-132); the
//      checksum is hashed on the device beside the key, the host only encodes.
//
// The re-derivation shares nothing with the search but the curve: a fixed-base
// comb over B in the 8 x 32 field of fe.hpp (64 four-bit windows, the 16
// multiples j * 16^i * B of each with Z = 1, built on the host), where the
// search walks a chain of additions of 8G in the 10-limb field.  Scalars are
// secret, but this runs on the machine that already holds every secret it
// derives, so the table index is not hidden.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"
#include "keccak.hpp"
#include "og_host.hpp"

using namespace og;

namespace {

constexpr int kCombWindows = 64;  // four bits each
constexpr int kCombWords = 24;    // y+x, y-x, 2dxy: 8 words each

// out: per hit the packed public key (8 words) and its address checksum (one
// word), so the host only has to base32-encode.
__global__ __launch_bounds__(256) void og_base_mult_kernel(const uint32_t *scalars,
                                                           const uint32_t *comb, uint32_t *out,
                                                           uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) s[l] = scalars[(size_t)i * 8 + l];
  Point p = pt_identity();
  for (int w = 0; w < kCombWindows; ++w) {
    const uint32_t j = (s[w >> 3] >> (4 * (w & 7))) & 15u;
    const uint32_t *e = comb + ((size_t)w * 16 + j) * kCombWords;
    Cached1 q;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      q.ypx.v[l] = e[l];
      q.ymx.v[l] = e[8 + l];
      q.t2d.v[l] = e[16 + l];
    }
    p = pt_add1(p, q);
  }
  const Fe zinv = fe_invert(p.Z);
  const Fe y = fe_canon(fe_mul(p.Y, zinv));
  const Fe x = fe_canon(fe_mul(p.X, zinv));
  uint32_t pk[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) pk[l] = y.v[l];
  pk[7] |= (x.v[0] & 1u) << 31;
#pragma unroll
  for (int l = 0; l < 8; ++l) out[(size_t)i * 9 + l] = pk[l];
  out[(size_t)i * 9 + 8] = onion_checksum(pk);
}

// base32 of the 35 address bytes (src/base32.rs:31-88): 56 symbols.
void address_text(const uint8_t pk[32], uint32_t chk, char out[56]) {
  static const char kAlphabet[] = "abcdefghijklmnopqrstuvwxyz234567";
  uint8_t b[35];
  std::memcpy(b, pk, 32);
  b[32] = (uint8_t)chk;
  b[33] = (uint8_t)(chk >> 8);
  b[34] = 0x03;
  for (int m = 0; m < 56; ++m) {
    const int bit = 5 * m, byte = bit >> 3, sh = bit & 7;
    const unsigned two = ((unsigned)b[byte] << 8) | (byte + 1 < 35 ? b[byte + 1] : 0u);
    out[m] = kAlphabet[(two >> (11 - sh)) & 31u];
  }
}

void put_err(char *err, size_t cap, const std::string &msg) {
  if (err && cap) {
    size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

// j * 16^w * B for w < 64, j < 16, affine with Z = 1 (j = 0: the identity).
const std::vector<uint32_t> &comb_table() {
  static std::vector<uint32_t> t;
  static std::once_flag once;
  std::call_once(once, [] {
    const HostCurve hc;
    t.assign((size_t)kCombWindows * 16 * kCombWords, 0);
    Point base = hc.base;  // 16^w * B
    for (int w = 0; w < kCombWindows; ++w) {
      const Cached bc = pt_to_cached(base, hc.two_d);
      Point acc = pt_identity();
      for (int j = 0; j < 16; ++j) {
        const Fe zinv = fe_invert(acc.Z);
        const Fe x = fe_mul(acc.X, zinv), y = fe_mul(acc.Y, zinv);
        const Fe ypx = fe_canon(fe_add(y, x)), ymx = fe_canon(fe_sub(y, x));
        const Fe t2d = fe_canon(fe_mul(hc.two_d, fe_mul(x, y)));
        uint32_t *e = t.data() + ((size_t)w * 16 + j) * kCombWords;
        for (int l = 0; l < 8; ++l) {
          e[l] = ypx.v[l];
          e[8 + l] = ymx.v[l];
          e[16 + l] = t2d.v[l];
        }
        acc = pt_add(acc, bc);  // (j + 1) * 16^w * B
      }
      base = acc;  // 16 * 16^w * B
    }
  });
  return t;
}

}  // namespace

extern "C" {

int og_hip_finish_hits(int ordinal, const uint8_t seed[32], const og_survivor *hits, size_t n,
                       og_key_record *out, char *err, size_t errcap) {
  if (!seed || (n && (!hits || !out))) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  if (n == 0) return 0;
  if (n > (1u << 26)) {
    put_err(err, errcap, "at most 2^26 hits per call");
    return -1;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || ordinal < 0 || ordinal >= count) {
    put_err(err, errcap, "no such HIP device");
    return -1;
  }
  uint8_t base[64];
  og_key_expanded_secret(seed, base);
  std::vector<uint32_t> scalars(n * 8);
  for (size_t i = 0; i < n; ++i) {
    og_key_record &r = out[i];
    std::memset(&r, 0, sizeof r);
    r.offset = hits[i].offset;
    // add_to_scalar on the clamped left half; the right half is carried over
    std::memcpy(r.expanded_secret, base, 64);
    unsigned __int128 carry = hits[i].offset;
    for (int b = 0; b < 32 && carry; ++b) {
      carry += r.expanded_secret[b];
      r.expanded_secret[b] = (uint8_t)carry;
      carry >>= 8;
    }
    const uint8_t lo = r.expanded_secret[0], hi = r.expanded_secret[31];
    r.status = ((lo & 248) == lo && ((hi & 63) | 64) == hi) ? OG_HIT_OK : OG_HIT_CLAMPING;
    std::memcpy(&scalars[i * 8], r.expanded_secret, 32);
  }
  const std::vector<uint32_t> &comb = comb_table();
  uint32_t *d_scalars = nullptr, *d_comb = nullptr, *d_out = nullptr;
  std::vector<uint32_t> pks(n * 9);
  hipError_t e = hipSetDevice(ordinal);
  if (e == hipSuccess) e = hipMalloc(&d_scalars, n * 32);
  if (e == hipSuccess) e = hipMalloc(&d_out, n * 36);
  if (e == hipSuccess) e = hipMalloc(&d_comb, comb.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(d_scalars, scalars.data(), n * 32, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_comb, comb.data(), comb.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(og_base_mult_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       d_scalars, d_comb, d_out, (uint32_t)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(pks.data(), d_out, n * 36, hipMemcpyDeviceToHost);
  (void)hipFree(d_scalars);
  (void)hipFree(d_out);
  (void)hipFree(d_comb);
  if (e != hipSuccess) {
    put_err(err, errcap, std::string("public-key re-derivation: ") + hipGetErrorString(e));
    return -1;
  }
  for (size_t i = 0; i < n; ++i) {
    og_key_record &r = out[i];
    if (r.status != OG_HIT_OK) continue;
    std::memcpy(r.public_key, &pks[i * 9], 32);
    // the survivor carries y with the sign bit clear: compare all 255 bits
    uint8_t y[32];
    std::memcpy(y, r.public_key, 32);
    y[31] &= 0x7f;
    if (std::memcmp(y, hits[i].packed, 32) != 0) {
      r.status = OG_HIT_MISMATCH;
      continue;
    }
    address_text(r.public_key, pks[i * 9 + 8], r.hostname);
    std::memcpy(r.hostname + 56, ".onion", 7);
  }
  return 0;
}

}  // extern "C"
