// Device engine for the candidate-key search: kernels, host engine, C ABI.
//
// Replaces, beneath the reference's gpu::Engine boundary (src/gpu/mod.rs:292-477),
// its vendor engine (src/gpu/cuda.rs:103-365) and kernels chain_steps /
// chain_reduce (cuda-kernel/src/lib.rs:52-236).  The kernel design is in
// og_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"
#include "og_filters.hpp"
#include "og_kernels.hpp"
#include "fe10.hpp"
#include "fe9.hpp"
#include "og_host.hpp"

using namespace og;

// ------------------------------------------------------------------ device

namespace {

typedef const __attribute__((address_space(4))) uint32_t *cptr;

__device__ __forceinline__ Fe ld_const(cptr p) {
  Fe r;
#pragma unroll
  for (int l = 0; l < 8; ++l) r.v[l] = p[l];
  return r;
}

__device__ __forceinline__ Fe ld_soa(const uint32_t *base, size_t plane, uint32_t T, uint32_t g) {
  Fe r;
#pragma unroll
  for (int l = 0; l < 8; ++l) r.v[l] = __builtin_nontemporal_load(base + (plane * 8 + l) * T + g);
  return r;
}

__device__ __forceinline__ void st_soa(uint32_t *base, size_t plane, uint32_t T, uint32_t g,
                                       const Fe &v) {
#pragma unroll
  for (int l = 0; l < 8; ++l) __builtin_nontemporal_store(v.v[l], base + (plane * 8 + l) * T + g);
}

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// Whether the canonical key y (sign clear) passes the key-side filters.
template <uint32_t MODE>
__device__ __forceinline__ bool key_test(const Fe &c, const DevFilterArgs &f) {
  if (MODE & kModePrefix) {
    uint32_t diff = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) diff |= (c.v[l] ^ f.want[l]) & f.mask[l];
    return diff == 0;
  }
  if (MODE & kModeBitmap) {
    const uint32_t head = bswap32(c.v[0]);
    const uint32_t slot = (uint32_t)(((uint64_t)head) >> (32 - f.bits));
    const bool in = (f.bitmap[slot >> 5] >> (slot & 31)) & 1u;
    if (!in || f.nbuckets == 0) return in;
    const uint64_t h = ((uint64_t)head << 32) | bswap32(c.v[1]);
    for (uint32_t b = 0; b < f.nbuckets; ++b) {
      const uint64_t probe = h & f.bmask[b];
      const uint64_t *keys = f.keys + f.boff[b];
      uint32_t lo = 0, hi = f.bcnt[b];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < probe) lo = mid + 1;
        else hi = mid;
      }
      if (lo < f.bcnt[b] && keys[lo] == probe) return true;
    }
    return false;
  }
  return false;
}

// Symbol m (0..55) of the address held as nine big-endian words.
__device__ __forceinline__ uint32_t addr_symbol(const uint32_t w[9], uint32_t m) {
  const uint32_t bit = 5 * m, i = bit >> 5, sh = bit & 31;
  const uint32_t hi = w[i];
  const uint32_t lo = i + 1 < 9 ? w[i + 1] : 0u;
  const uint64_t both = ((uint64_t)hi << 32) | lo;
  return (uint32_t)(both >> (59 - sh)) & 31u;
}

__device__ __forceinline__ bool pattern_test(const uint32_t w[9], const DevFilterArgs &f) {
  for (uint32_t p = 0; p < f.npat; ++p) {
    const uint32_t *rec = f.pats + p * (4 + kAddressLen);
    const uint32_t len = rec[0], first = rec[1], last = rec[2];
    for (uint32_t s = first; s <= last; ++s) {
      bool ok = true;
      for (uint32_t i = 0; i < len && ok; ++i) ok = (rec[4 + i] >> addr_symbol(w, s + i)) & 1u;
      if (ok) return true;
    }
  }
  return false;
}

// Appends a wave's hits with one atomic per wave.
__device__ __forceinline__ void append(bool hit, uint64_t k, const Fe &y, uint32_t *counter,
                                       HitRec *hits, uint32_t cap) {
  const uint64_t mask = __ballot(hit);
  if (mask == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  if (hit) {
    const uint32_t idx = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
    if (idx < cap) {
      hits[idx].k = k;
#pragma unroll
      for (int l = 0; l < 8; ++l) hits[idx].y[l] = y.v[l];
    }
  }
}

// Full test of one candidate: y (unreduced), optional x numerator product.
template <uint32_t MODE>
__device__ __forceinline__ void test_candidate(const Fe &y, const Fe *x, uint64_t k,
                                               const SearchArgs &a) {
  const Fe c = fe_canon(y);
  bool hit = false;
  if (MODE & (kModePrefix | kModeBitmap)) hit = key_test<MODE>(c, a.f);
  if (MODE & kModePattern) {
    if (!hit) {
      const Fe cx = fe_canon(*x);
      uint32_t pk[8];
#pragma unroll
      for (int l = 0; l < 8; ++l) pk[l] = c.v[l];
      pk[7] |= (cx.v[0] & 1u) << 31;
      const uint32_t chk = onion_checksum(pk);
      uint32_t w[9];
#pragma unroll
      for (int l = 0; l < 8; ++l) w[l] = bswap32(pk[l]);
      w[8] = ((chk & 0xffu) << 24) | (((chk >> 8) & 0xffu) << 16) | (0x03u << 8);
      hit = pattern_test(w, a.f);
    }
  }
  append(hit, k, c, a.counter, a.hits, a.cap);
}

template <uint32_t MODE>
__global__ __launch_bounds__(256) void og_search_kernel(SearchArgs a) {
  constexpr bool kX = (MODE & kModePattern) != 0;
  const uint32_t T = a.threads;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t H = a.half;
  const uint64_t S = 2ull * H + 1;
  cptr table = (cptr)a.table;

  Point P;
  P.X = ld_soa(a.state, 0, T, g);
  P.Y = ld_soa(a.state, 1, T, g);
  P.Z = ld_soa(a.state, 2, T, g);
  P.T = ld_soa(a.state, 3, T, g);
  Cached1 st;
  for (int l = 0; l < 8; ++l) {
    st.ypx.v[l] = a.stride[l];
    st.ymx.v[l] = a.stride[8 + l];
    st.t2d.v[l] = a.stride[16 + l];
  }
  // x numerators use a second scratch region after the y numerators.
  uint32_t *xs = a.scratch + (size_t)2 * H * 8 * T;

  for (uint32_t r = 0; r < a.rounds; ++r) {
    const uint64_t w = (uint64_t)r * T + g;
    const uint64_t kc = a.k_base + w * S + H;
    Fe D = fe_one();
    for (uint32_t m = 0; m < H; ++m) {
      const Fe xq = ld_const(table + m * 24);
      const Fe yq = ld_const(table + m * 24 + 8);
      const Fe dq = ld_const(table + m * 24 + 16);
      const Fe A = fe_mul(P.Y, yq);
      const Fe B = fe_mul(P.X, xq);
      const Fe C = fe_mul(P.T, dq);
      const Fe dp = fe_sub(P.Z, C);
      const Fe dm = fe_add(P.Z, C);
      const Fe D0 = D;
      st_soa(a.scratch, 2 * m, T, g, fe_mul(fe_add(A, B), D));
      D = fe_mul(D, dp);
      st_soa(a.scratch, 2 * m + 1, T, g, fe_mul(fe_sub(A, B), D));
      if (kX) {
        const Fe XY = fe_mul(P.X, yq);
        const Fe YX = fe_mul(P.Y, xq);
        // x(P+Q) = (X yq + Y xq)/dm shares dm's prefix D (= D0*dp);
        // x(P-Q) = (X yq - Y xq)/dp shares dp's prefix D0.
        st_soa(xs, 2 * m, T, g, fe_mul(fe_add(XY, YX), D));
        st_soa(xs, 2 * m + 1, T, g, fe_mul(fe_sub(XY, YX), D0));
      }
      D = fe_mul(D, dm);
    }
    const Fe sc = fe_mul(P.Y, D);
    const Fe tc = kX ? fe_mul(P.X, D) : D;
    D = fe_mul(D, P.Z);
    Fe inv = fe_invert(D);
    {
      const Fe y = fe_mul(sc, inv);
      const Fe x = kX ? fe_mul(tc, inv) : y;
      test_candidate<MODE>(y, &x, kc, a);
    }
    inv = fe_mul(inv, P.Z);
    for (uint32_t m = H; m-- > 0;) {
      const Fe dq = ld_const(table + m * 24 + 16);
      const Fe C = fe_mul(P.T, dq);
      const Fe dp = fe_sub(P.Z, C);
      const Fe dm = fe_add(P.Z, C);
      // entry: inv = 1/(D0 dp dm).  y(P-Q) = s_m inv; x(P+Q) = t_p inv;
      // after inv *= dm: y(P+Q) = s_p inv, x(P-Q) = t_m inv.
      const Fe ym = fe_mul(ld_soa(a.scratch, 2 * m + 1, T, g), inv);
      const Fe xp = kX ? fe_mul(ld_soa(xs, 2 * m, T, g), inv) : ym;
      inv = fe_mul(inv, dm);
      const Fe xm = kX ? fe_mul(ld_soa(xs, 2 * m + 1, T, g), inv) : ym;
      test_candidate<MODE>(ym, &xm, kc - (m + 1), a);
      const Fe yp = fe_mul(ld_soa(a.scratch, 2 * m, T, g), inv);
      test_candidate<MODE>(yp, &xp, kc + (m + 1), a);
      inv = fe_mul(inv, dp);
    }
    P = pt_add1(P, st);
  }
  st_soa(a.state, 0, T, g, P.X);
  st_soa(a.state, 1, T, g, P.Y);
  st_soa(a.state, 2, T, g, P.Z);
  st_soa(a.state, 3, T, g, P.T);
}

__global__ __launch_bounds__(256) void og_setup_kernel(SetupArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.threads) return;
  Cached eight, start;
  Point p0;
  for (int l = 0; l < 8; ++l) {
    eight.ypx.v[l] = a.eight[l];
    eight.ymx.v[l] = a.eight[8 + l];
    eight.Z.v[l] = a.eight[16 + l];
    eight.t2d.v[l] = a.eight[24 + l];
    p0.X.v[l] = a.start[l];
    p0.Y.v[l] = a.start[8 + l];
    p0.Z.v[l] = a.start[16 + l];
    p0.T.v[l] = a.start[24 + l];
  }
  Fe two_d;
  for (int l = 0; l < 8; ++l) two_d.v[l] = a.two_d[l];
  start = pt_to_cached(p0, two_d);
  const uint64_t S = 2ull * a.half + 1;
  const uint64_t m = (uint64_t)g * S + a.half;
  Point p = pt_add(pt_mul_u64(m, eight), start);
  st_soa(a.state, 0, a.threads, g, p.X);
  st_soa(a.state, 1, a.threads, g, p.Y);
  st_soa(a.state, 2, a.threads, g, p.Z);
  st_soa(a.state, 3, a.threads, g, p.T);
}

// Elementwise field operations for the parity tests of the arithmetic itself:
// op 0 mul, 1 sq, 2 add, 3 sub, 4 invert, 5 canonical; outputs canonical.
__global__ __launch_bounds__(256) void og_field_kernel(int op, const uint32_t *a, const uint32_t *b,
                                                       uint32_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fe x, y;
  for (int l = 0; l < 8; ++l) {
    x.v[l] = a[(size_t)i * 8 + l];
    y.v[l] = b[(size_t)i * 8 + l];
  }
  Fe r;
  switch (op) {
    case 0: r = fe_mul(x, y); break;
    case 1: r = fe_sq(x); break;
    case 2: r = fe_add(x, y); break;
    case 3: r = fe_sub(x, y); break;
    case 4: r = fe_invert(x); break;
    default: r = x; break;
  }
  r = fe_canon(r);
  for (int l = 0; l < 8; ++l) out[(size_t)i * 8 + l] = r.v[l];
}

// Radix bench: each lane runs two independent chains of `iters` dependent
// multiplications x <- x*y in the chosen representation (8: 8 x 32 saturated,
// 10: 10 x 25.5), then writes the canonical results.  Inputs are a fixed
// function of the lane (radix_input), so the host can recompute the answer.
__host__ __device__ inline uint32_t radix_input(uint32_t lane, uint32_t which, uint32_t l) {
  uint32_t x = lane * 0x9E3779B9u + which * 0x85EBCA6Bu + l * 0xC2B2AE35u + 0x27220A95u;
  x ^= x >> 15;
  x *= 0x2C1B3C6Du;
  x ^= x >> 12;
  x *= 0x297A2D39u;
  x ^= x >> 15;
  return l == 7 ? (x & 0x7fffffffu) : x;
}

template <int RADIX>
__global__ __launch_bounds__(256) void og_radix_kernel(uint32_t iters, uint32_t *out, uint32_t n) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  Fe a, b, y;
  for (int l = 0; l < 8; ++l) {
    a.v[l] = radix_input(g, 0, l);
    b.v[l] = radix_input(g, 1, l);
    y.v[l] = radix_input(g, 2, l);
  }
  if (RADIX == 8) {
    for (uint32_t i = 0; i < iters; ++i) {
      a = fe_mul(a, y);
      b = fe_mul(b, y);
    }
  } else if (RADIX == 81) {
    for (uint32_t i = 0; i < iters; ++i) {
      a = fe_mul_madc(a, y);
      b = fe_mul_madc(b, y);
    }
  } else if (RADIX == 9) {
    Fe9 a9 = fe9_from(a), b9 = fe9_from(b), y9 = fe9_from(y);
    for (uint32_t i = 0; i < iters; ++i) {
      a9 = fe9_mul(a9, y9);
      b9 = fe9_mul(b9, y9);
    }
    a = fe9_to(a9);
    b = fe9_to(b9);
  } else {
    Fe10 a10 = fe10_from(a), b10 = fe10_from(b), y10 = fe10_from(y);
    for (uint32_t i = 0; i < iters; ++i) {
      a10 = fe10_mul(a10, y10);
      b10 = fe10_mul(b10, y10);
    }
    a = fe10_to(a10);
    b = fe10_to(b10);
  }
  a = fe_canon(a);
  b = fe_canon(b);
  for (int l = 0; l < 8; ++l) {
    out[(size_t)g * 16 + l] = a.v[l];
    out[(size_t)g * 16 + 8 + l] = b.v[l];
  }
}

}  // namespace

// -------------------------------------------------------------------- host

struct og_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t mode = 0;
  uint32_t threads = 0, half = 0, rounds = 1, block = 256, cap = 0;
  uint64_t launches = 0;
  uint32_t *d_state = nullptr, *d_table = nullptr, *d_scratch = nullptr, *d_bitmap = nullptr;
  uint64_t *d_keys = nullptr;
  uint32_t *d_pats = nullptr;
  uint32_t *d_counter = nullptr;
  HitRec *d_hits = nullptr;
  uint32_t *h_counter = nullptr;
  HitRec *h_hits = nullptr;
  SearchArgs args;
  float last_ms = 0;
};

namespace {

void put_err(char *err, size_t cap, const std::string &msg) {
  if (err && cap) {
    size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

#define OG_TRY(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      put_err(err, errcap, std::string(#expr) + ": " + hipGetErrorString(e_));       \
      return -1;                                                                     \
    }                                                                                \
  } while (0)

void free_engine(og_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  (void)hipFree(e->d_state);
  (void)hipFree(e->d_table);
  (void)hipFree(e->d_scratch);
  (void)hipFree(e->d_bitmap);
  (void)hipFree(e->d_keys);
  (void)hipFree(e->d_pats);
  (void)hipFree(e->d_counter);
  (void)hipFree(e->d_hits);
  (void)hipHostFree(e->h_counter);
  (void)hipHostFree(e->h_hits);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

template <uint32_t MODE>
hipError_t launch_search(og_engine *e) {
  const uint32_t grid = e->threads / e->block;
  hipLaunchKernelGGL(og_search_kernel<MODE>, dim3(grid), dim3(e->block), 0, e->stream, e->args);
  return hipGetLastError();
}

hipError_t launch_mode(og_engine *e) {
  switch (e->mode) {
    case kModePrefix: return launch_search<kModePrefix>(e);
    case kModeBitmap: return launch_search<kModeBitmap>(e);
    case kModePattern: return launch_search<kModePattern>(e);
    case kModePrefix | kModePattern: return launch_search<kModePrefix | kModePattern>(e);
    case kModeBitmap | kModePattern: return launch_search<kModeBitmap | kModePattern>(e);
    default: return hipErrorInvalidValue;
  }
}

int create_engine(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                  const FilterSet *fs, const uint64_t *bitmap, size_t bitmap_words,
                  uint32_t bitmap_bits, og_engine **out, char *err, size_t errcap) {
  if (!out || !start_xyzt) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    put_err(err, errcap, "no HIP device is visible");
    return -1;
  }
  if (ordinal < 0 || ordinal >= count) {
    put_err(err, errcap, "device ordinal out of range");
    return -1;
  }
  og_engine_opts o = opts ? *opts : og_engine_opts{};
  if (!o.threads) o.threads = 1u << 17;
  if (!o.half) o.half = 256;
  if (!o.rounds) o.rounds = 1;
  if (o.threads % 256 || o.threads > (1u << 22)) {
    put_err(err, errcap, "threads must be a positive multiple of 256, at most 2^22");
    return -1;
  }
  if (o.half > 4096) {
    put_err(err, errcap, "half must be at most 4096");
    return -1;
  }
  std::unique_ptr<og_engine, void (*)(og_engine *)> e(new og_engine(), free_engine);
  e->device = ordinal;
  e->threads = o.threads;
  e->half = o.half;
  e->rounds = o.rounds;
  OG_TRY(hipSetDevice(ordinal));
  OG_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  OG_TRY(hipEventCreate(&e->ev0));
  OG_TRY(hipEventCreate(&e->ev1));

  // ---- filters -> device tables
  SearchArgs &a = e->args;
  std::memset(&a, 0, sizeof a);
  std::vector<uint32_t> bm32;
  uint32_t bits = 0;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> pats;
  if (fs) {
    const bool one_prefix = fs->fast.size() == 1;
    if (one_prefix) {
      const Filter &f = fs->filters[fs->fast[0]];
      e->mode |= kModePrefix;
      for (int i = 0; i < f.nbytes; ++i) {
        const uint8_t m = i == f.nbytes - 1 ? f.mask : 0xff;
        a.f.want[i / 4] |= (uint32_t)f.bytes[i] << (8 * (i % 4));
        a.f.mask[i / 4] |= (uint32_t)m << (8 * (i % 4));
      }
    } else if (!fs->fast.empty()) {
      e->mode |= kModeBitmap;
      std::vector<uint64_t> w = fs->index_bits ? fs->words : bitmap_for(*fs, 24);
      bits = fs->index_bits ? fs->index_bits : 24;
      for (uint64_t v : w) {
        bm32.push_back((uint32_t)v);
        bm32.push_back((uint32_t)(v >> 32));
      }
      if (fs->buckets.size() > (size_t)kMaxBuckets) {
        put_err(err, errcap, "more distinct prefix lengths than the device table holds");
        return -1;
      }
      for (const Bucket &b : fs->buckets) {
        a.f.boff[a.f.nbuckets] = (uint32_t)keys.size();
        a.f.bcnt[a.f.nbuckets] = (uint32_t)b.keys.size();
        a.f.bmask[a.f.nbuckets] = b.mask;
        a.f.nbuckets++;
        keys.insert(keys.end(), b.keys.begin(), b.keys.end());
      }
    }
    if (!fs->general.empty()) {
      if (fs->general.size() > (size_t)kMaxPatterns) {
        put_err(err, errcap, "at most 8 general (suffix, contains, class) filters per search");
        return -1;
      }
      e->mode |= kModePattern;
      for (uint32_t i : fs->general) {
        DevPattern p = dev_pattern(fs->filters[i]);
        pats.push_back(p.len);
        pats.push_back(p.first);
        pats.push_back(p.last);
        pats.push_back(p.key_last);
        for (int s = 0; s < kAddressLen; ++s) pats.push_back(p.sets[s]);
      }
      a.f.npat = (uint32_t)fs->general.size();
    }
    if (e->mode == 0) {
      put_err(err, errcap, "the filter set is empty");
      return -1;
    }
  } else {
    if (!bitmap || bitmap_bits < 8 || bitmap_bits > 32 ||
        bitmap_words < (((size_t)1 << bitmap_bits) + 63) / 64) {
      put_err(err, errcap, "a bitmap of 8..32 bits with all its words is required");
      return -1;
    }
    e->mode = kModeBitmap;
    bits = bitmap_bits;
    const size_t nw = (((size_t)1 << bitmap_bits) + 63) / 64;
    for (size_t i = 0; i < nw; ++i) {
      bm32.push_back((uint32_t)bitmap[i]);
      bm32.push_back((uint32_t)(bitmap[i] >> 32));
    }
  }
  a.f.bits = bits;

  // ---- constants
  const uint64_t S = 2ull * o.half + 1;
  HostCurve hc;
  std::vector<uint32_t> table = hc.offset_table(o.half);
  Cached1 stride = hc.multiple_of_eight_cached1((uint64_t)o.threads * S);
  for (int l = 0; l < 8; ++l) {
    a.stride[l] = stride.ypx.v[l];
    a.stride[8 + l] = stride.ymx.v[l];
    a.stride[16 + l] = stride.t2d.v[l];
  }

  // ---- buffers
  const bool wantx = (e->mode & kModePattern) != 0;
  const size_t scratch_words = (size_t)2 * o.half * 8 * o.threads * (wantx ? 2 : 1);
  e->cap = o.hit_cap ? o.hit_cap : og_hit_room((uint64_t)o.threads * S * o.rounds, bits ? bits : 24);
  OG_TRY(hipMalloc(&e->d_state, (size_t)32 * o.threads * 4));
  OG_TRY(hipMalloc(&e->d_table, std::max<size_t>(table.size(), 24) * 4));
  OG_TRY(hipMalloc(&e->d_scratch, std::max<size_t>(scratch_words, 1) * 4));
  OG_TRY(hipMalloc(&e->d_counter, 4));
  OG_TRY(hipMalloc(&e->d_hits, (size_t)e->cap * sizeof(HitRec)));
  OG_TRY(hipHostMalloc(&e->h_counter, 4, hipHostMallocDefault));
  OG_TRY(hipHostMalloc(&e->h_hits, (size_t)e->cap * sizeof(HitRec), hipHostMallocDefault));
  if (!table.empty())
    OG_TRY(hipMemcpyAsync(e->d_table, table.data(), table.size() * 4, hipMemcpyHostToDevice, e->stream));
  if (!bm32.empty()) {
    OG_TRY(hipMalloc(&e->d_bitmap, bm32.size() * 4));
    OG_TRY(hipMemcpyAsync(e->d_bitmap, bm32.data(), bm32.size() * 4, hipMemcpyHostToDevice, e->stream));
  }
  if (!keys.empty()) {
    OG_TRY(hipMalloc(&e->d_keys, keys.size() * 8));
    OG_TRY(hipMemcpyAsync(e->d_keys, keys.data(), keys.size() * 8, hipMemcpyHostToDevice, e->stream));
  }
  if (!pats.empty()) {
    OG_TRY(hipMalloc(&e->d_pats, pats.size() * 4));
    OG_TRY(hipMemcpyAsync(e->d_pats, pats.data(), pats.size() * 4, hipMemcpyHostToDevice, e->stream));
  }
  a.state = e->d_state;
  a.table = e->d_table;
  a.scratch = e->d_scratch;
  a.threads = o.threads;
  a.half = o.half;
  a.rounds = o.rounds;
  a.f.bitmap = e->d_bitmap;
  a.f.keys = e->d_keys;
  a.f.pats = e->d_pats;
  a.counter = e->d_counter;
  a.hits = e->d_hits;
  a.cap = e->cap;

  // ---- per-chain start points on the device
  SetupArgs sa;
  std::memset(&sa, 0, sizeof sa);
  sa.state = e->d_state;
  sa.threads = o.threads;
  sa.half = o.half;
  for (int c = 0; c < 4; ++c) {
    Fe v = fe_from_bytes(start_xyzt + 32 * c);
    for (int l = 0; l < 8; ++l) sa.start[8 * c + l] = v.v[l];
  }
  Cached eight = hc.eight_cached();
  for (int l = 0; l < 8; ++l) {
    sa.eight[l] = eight.ypx.v[l];
    sa.eight[8 + l] = eight.ymx.v[l];
    sa.eight[16 + l] = eight.Z.v[l];
    sa.eight[24 + l] = eight.t2d.v[l];
    sa.two_d[l] = hc.two_d.v[l];
  }
  hipLaunchKernelGGL(og_setup_kernel, dim3(o.threads / 256), dim3(256), 0, e->stream, sa);
  OG_TRY(hipGetLastError());
  OG_TRY(hipStreamSynchronize(e->stream));
  *out = e.release();
  return 0;
}

}  // namespace

extern "C" {

int og_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int og_hip_device_info(int ordinal, char *name, size_t namecap, uint64_t *memory,
                       int *compute_units) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, ordinal) != hipSuccess) return -1;
  if (name && namecap) {
    std::snprintf(name, namecap, "%s (%s)", p.name, p.gcnArchName);
  }
  if (memory) *memory = (uint64_t)p.totalGlobalMem;
  if (compute_units) *compute_units = p.multiProcessorCount;
  return 0;
}

uint32_t og_hit_room(uint64_t batch, uint32_t bitmap_bits) {
  // src/gpu/mod.rs:306-312: four times the expected survivors plus 1024,
  // clamped to [4096, 2^18].
  const uint64_t expected = batch >> std::min<uint32_t>(bitmap_bits, 63);
  uint64_t room = expected > (UINT64_MAX - 1024) / 4 ? UINT64_MAX : expected * 4 + 1024;
  room = std::min<uint64_t>(std::max<uint64_t>(room, 4096), 1u << 18);
  return (uint32_t)room;
}

int og_hip_engine_create(int ordinal, const uint8_t start_xyzt[128], uint32_t threads,
                         const uint64_t *bitmap, size_t bitmap_words, uint32_t bitmap_bits,
                         og_engine **out, char *err, size_t errcap) {
  og_engine_opts o{};
  o.threads = threads;
  return create_engine(ordinal, start_xyzt, &o, nullptr, bitmap, bitmap_words, bitmap_bits, out,
                       err, errcap);
}

int og_hip_engine_create_ex(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                            const og_filterset *filters, og_engine **out, char *err,
                            size_t errcap) {
  if (!filters) {
    put_err(err, errcap, "null filter set");
    return -1;
  }
  return create_engine(ordinal, start_xyzt, opts, filterset_of(filters), nullptr, 0, 0, out, err,
                       errcap);
}

uint32_t og_hip_engine_threads(const og_engine *e) { return e ? e->threads : 0; }

uint64_t og_hip_engine_batch(const og_engine *e) {
  return e ? (uint64_t)e->threads * (2ull * e->half + 1) * e->rounds : 0;
}

double og_hip_engine_last_ms(const og_engine *e) { return e ? e->last_ms : 0.0; }

int og_hip_engine_advance(og_engine *e, og_survivor *out, uint32_t cap, uint32_t *n, char *err,
                          size_t errcap) {
  if (!e || !n) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  *n = 0;
  OG_TRY(hipSetDevice(e->device));
  e->args.k_base = e->launches * og_hip_engine_batch(e);
  OG_TRY(hipMemsetAsync(e->d_counter, 0, 4, e->stream));
  OG_TRY(hipEventRecord(e->ev0, e->stream));
  OG_TRY(launch_mode(e));
  OG_TRY(hipEventRecord(e->ev1, e->stream));
  OG_TRY(hipMemcpyAsync(e->h_counter, e->d_counter, 4, hipMemcpyDeviceToHost, e->stream));
  OG_TRY(hipStreamSynchronize(e->stream));
  OG_TRY(hipEventElapsedTime(&e->last_ms, e->ev0, e->ev1));
  e->launches++;
  const uint32_t found = *e->h_counter;
  const uint32_t kept = std::min(found, e->cap);
  if (kept) {
    OG_TRY(hipMemcpyAsync(e->h_hits, e->d_hits, (size_t)kept * sizeof(HitRec),
                          hipMemcpyDeviceToHost, e->stream));
    OG_TRY(hipStreamSynchronize(e->stream));
  }
  if (found > e->cap) {
    put_err(err, errcap,
            "one launch produced " + std::to_string(found) + " survivors and only " +
                std::to_string(e->cap) + " fit; the index is too narrow for what it is filtering");
    return -2;
  }
  if (kept > cap) {
    put_err(err, errcap, "the output buffer holds " + std::to_string(cap) + " survivors, " +
                             std::to_string(kept) + " were found");
    return -3;
  }
  for (uint32_t i = 0; i < kept; ++i) {
    out[i].offset = e->h_hits[i].k * 8;
    for (int l = 0; l < 8; ++l)
      for (int b = 0; b < 4; ++b) out[i].packed[4 * l + b] = (uint8_t)(e->h_hits[i].y[l] >> (8 * b));
  }
  *n = kept;
  return 0;
}

void og_hip_engine_destroy(og_engine *e) { free_engine(e); }

int og_hip_radix_bench(int ordinal, int radix, uint32_t threads, uint32_t iters,
                       double *muls_per_s, uint8_t *results, char *err, size_t errcap) {
  if ((radix != 8 && radix != 81 && radix != 9 && radix != 10) || threads == 0 || threads % 256 || threads > (1u << 22)) {
    put_err(err, errcap, "radix must be 8, 81, 9 or 10 and threads a multiple of 256 up to 2^22");
    return -1;
  }
  OG_TRY(hipSetDevice(ordinal));
  uint32_t *d = nullptr;
  OG_TRY(hipMalloc(&d, (size_t)threads * 64));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&](uint32_t it) {
    if (radix == 8)
      hipLaunchKernelGGL(og_radix_kernel<8>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
    else if (radix == 81)
      hipLaunchKernelGGL(og_radix_kernel<81>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
    else if (radix == 9)
      hipLaunchKernelGGL(og_radix_kernel<9>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
    else
      hipLaunchKernelGGL(og_radix_kernel<10>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
  };
  launch(iters < 16 ? iters : 16);  // warm-up
  (void)hipEventRecord(e0, 0);
  launch(iters);
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess && results) e = hipMemcpy(results, d, (size_t)threads * 64, hipMemcpyDeviceToHost);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d);
  if (e != hipSuccess) {
    put_err(err, errcap, std::string("radix bench: ") + hipGetErrorString(e));
    return -1;
  }
  if (muls_per_s) *muls_per_s = 2.0 * threads * (double)iters / (ms * 1e-3);
  return 0;
}

int og_hip_field_op(int ordinal, int op, const uint8_t *a, const uint8_t *b, uint8_t *out,
                    size_t n, char *err, size_t errcap) {
  if (op < 0 || op > 5 || !a || !b || !out) {
    put_err(err, errcap, "bad field operation or null buffer");
    return -1;
  }
  if (n == 0) return 0;
  if (n > (1u << 24)) {
    put_err(err, errcap, "at most 2^24 elements per call");
    return -1;
  }
  OG_TRY(hipSetDevice(ordinal));
  uint32_t *da = nullptr, *db = nullptr, *dout = nullptr;
  const size_t bytes = n * 32;
  hipError_t e1 = hipMalloc(&da, bytes), e2 = hipMalloc(&db, bytes), e3 = hipMalloc(&dout, bytes);
  int rc = 0;
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    put_err(err, errcap, "hipMalloc failed for the field test buffers");
    rc = -1;
  } else {
    hipError_t e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(og_field_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op,
                         da, db, dout, (uint32_t)n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      put_err(err, errcap, std::string("field kernel: ") + hipGetErrorString(e));
      rc = -1;
    }
  }
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dout);
  return rc;
}

}  // extern "C"
