// Device engine for the candidate-key search: kernels, host engine, C ABI.
//
// Replaces, beneath the reference's gpu::Engine boundary (src/gpu/mod.rs:292-477),
// its vendor engine (src/gpu/cuda.rs:103-365) and kernels chain_steps /
// chain_reduce (cuda-kernel/src/lib.rs:52-236).  The kernel design is in
// og_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <ctime>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"
#include "fe51.hpp"
#include "fe9.hpp"
#include "og_devfilter.hpp"
#include "og_filters.hpp"
#include "og_host.hpp"
#include "og_search.hpp"

using namespace og;

// ------------------------------------------------------------------ device

namespace {

template <uint32_t MODE>
__global__ __launch_bounds__(256) OG_SEARCH_ATTR(MODE) void og_search_kernel(SearchArgs a) {
  const uint32_t T = a.threads;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  Point10 P = ld_point(a.state, T, g);
  Cached10 st;
#pragma unroll
  for (int l = 0; l < 10; ++l) {
    st.ypx.v[l] = a.stride[l];
    st.ymx.v[l] = a.stride[10 + l];
    st.t2d.v[l] = a.stride[20 + l];
  }
  __shared__ ScanBank banks[(MODE & kModePattern) ? kMaxBanks : 1];
  // the first automaton, or a dictionary's window prefilter maps
  __shared__ uint32_t pattern_lds[(MODE & kModePattern) ? kPatternLdsWords : 1];
  uint16_t *dfa_lds = reinterpret_cast<uint16_t *>(pattern_lds);
  if (MODE & kModePattern) load_filter(banks, dfa_lds, a.f, true);
  // the key forms' bitmap probe reads this workgroup's LDS copy of the map,
  // or of its coarse fold, before (or instead of) the map in L2
  constexpr bool kCoarse = (MODE & kModeBitmap) && !(MODE & kModePattern);
  __shared__ uint32_t coarse_lds[kCoarse ? kCoarseWords : 1];
  const uint32_t *coarse = nullptr;
  if (kCoarse && a.f.coarse_bits) {
    const uint32_t words = (1u << a.f.coarse_bits) >> 5;
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) coarse_lds[i] = a.f.coarse[i];
    __syncthreads();
    coarse = coarse_lds;
  }
  HitSink sink{a.counter, a.hits, a.cap, 0, nullptr, 0};
  search_chain<MODE>(P, st, (cptr)a.table, a.scratch, T, g, a.half, a.rounds, T, g, a.k_base, a.f,
                     banks, dfa_lds, a.parked, sink, coarse);
  st_point(a.state, T, g, P);
}

// The deferred address stage of a one-round pattern launch: per chain, the
// candidates the key stage parked (SHA3-256, the exact automaton pass, and
// with kSign the sign of x first).  Without the search loop's registers it
// runs at four waves per SIMD instead of two; the form that derives the signs
// (products, an inversion per chunk) is its own instantiation, held to two
// waves so that it does not spill, and the plain one keeps its registers.
// The LDS holds the automaton banks (and with kMaps the window prefilter's
// maps), never the expressions' automata, which the address stage walks in
// global memory: without maps the registers alone bound the waves per SIMD.
// kMaps: the set's window prefilter serves the address stage; its two maps
// (64 KiB) take the LDS, in 1024-thread workgroups when the launch allows it
// (one per CU: four waves per SIMD).  Two 640-thread workgroups per CU, five
// waves per SIMD, measured slower: 367 against 314 ms for 1,000 five-symbol
// words, whose 104 LDS reads per candidate meet more bank conflicts.
template <bool kSign, bool kMaps>
constexpr uint32_t kAddressBlock = kMaps ? 1024u : 256u;

template <bool kSign, bool kMaps>
__global__ __launch_bounds__((kAddressBlock<kSign, kMaps>))
    __attribute__((amdgpu_waves_per_eu(kSign ? 2 : 4, 8))) void og_address_kernel(SearchArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ ScanBank banks[kMaxBanks];
  __shared__ uint32_t lds_words[kMaps ? kPatternLdsWords : 1];
  uint16_t *dfa_lds = reinterpret_cast<uint16_t *>(lds_words);
  load_filter(banks, dfa_lds, a.f, kMaps, false);
  if (g >= a.threads) return;  // the last workgroup of a launch of fewer chains
  HitSink sink{a.counter, a.hits, a.cap, 0, nullptr, 0};
  const uint32_t top = kPlanesPerPair * a.half;
  const lptr32 maps = kMaps ? (lptr32)lds_words : (lptr32) nullptr;
  if (kSign)
    drain_deferred(a.scratch, top, a.half, a.threads, g, a.parked[g], a.k_base, (cptr)a.table,
                   a.f, banks, maps, sink);
  else
    drain<kModePattern>(a.scratch, top, a.threads, g, a.parked[g], a.f, banks, maps, sink);
}

__global__ __launch_bounds__(256) void og_setup_kernel(SetupArgs a) {
  const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.count) return;
  setup_chain(a, local);
}

// Elementwise field operations for the parity tests of the arithmetic itself:
// ops 0..5 in the 8 x 32 form and 10..15 in the 10-limb form the search runs
// in: mul, sq, add, sub, invert, canonical.  Outputs are canonical.
__host__ __device__ inline void field_op_one(int op, const uint32_t *a, const uint32_t *b,
                                             uint32_t *out) {
  Fe x, y;
  for (int l = 0; l < 8; ++l) {
    x.v[l] = a[l];
    y.v[l] = b[l];
  }
  Fe r;
  if (op == 20) {
    // the address checksum of a as a packed key (sign bit as given)
    for (int l = 0; l < 8; ++l) out[l] = 0;
    out[0] = onion_checksum(a);
    return;
  }
  if (op < 10) {
    switch (op) {
      case 0: r = fe_mul(x, y); break;
      case 1: r = fe_sq(x); break;
      case 2: r = fe_add(x, y); break;
      case 3: r = fe_sub(x, y); break;
      case 4: r = fe_invert(x); break;
      default: r = x; break;
    }
    r = fe_canon(r);
  } else {
    const Fe10 x10 = fe10_from(x), y10 = fe10_from(y);
    Fe10 r10;
    switch (op) {
      case 10: r10 = fe10_mul(x10, y10); break;
      case 11: r10 = fe10_sq(x10); break;
      case 12: r10 = fe10_add(x10, y10); break;
      case 13: r10 = fe10_sub(x10, y10); break;
      case 14: r10 = fe10_invert(x10); break;
      // loose inputs at the bounds the search reaches: (2x - y) * (x - y)
      case 16: r10 = fe10_mul_wide(fe10_sub(fe10_add(x10, x10), y10), fe10_sub(x10, y10)); break;
      // the loosest operands the search loops multiply: (x - y) times y, 3 x 1
      // units, with the 32-bit fold of the top carry
      case 25: r10 = fe10_mul(fe10_sub(x10, y10), y10); break;
      // the window inversion the search runs: the divstep recurrence
      case 26: r10 = fe10_invert_sg(x10); break;
      // a product's output canonicalised as the search kernels do it: x times
      // one with x mod 2^255 split as it stands (values in [p, 2^255) stay
      // above p), and x times y
      case 17: r = fe10_canon_product(fe10_mul(fe10_split(x), fe10_one())); break;
      case 18: r = fe10_canon_product(fe10_mul(x10, y10)); break;
      // bit 0 of the canonical product, as the search takes the sign of x
      // (x mod 2^255 split as it stands, as for op 17)
      case 19: r = fe_zero(); r.v[0] = fe10_parity_product(fe10_mul(fe10_split(x), y10)); break;
      // the key forms' probe word from four columns of (2x - y)(x - y), the
      // loosest operands of op 16, and whether it is exact (word 1)
      case 23: {
        bool exact;
        r = fe_zero();
        r.v[0] = fe10_mul_low32(fe10_sub(fe10_add(x10, x10), y10), fe10_sub(x10, y10), exact);
        r.v[1] = exact ? 1u : 0u;
        break;
      }
      // the same from three columns: bits 0..25 of the canonical product
      case 27: {
        bool exact;
        r = fe_zero();
        r.v[0] = fe10_mul_low26(fe10_sub(fe10_add(x10, x10), y10), fe10_sub(x10, y10), exact) &
                 kLow26;
        r.v[1] = exact ? 1u : 0u;
        break;
      }
      // bit 0 of the canonical x times y as the search takes the sign of x,
      // from three columns where they settle it (x mod 2^255 split as it
      // stands, as for op 19)
      case 24: r = fe_zero(); r.v[0] = fe10_parity_mul(fe10_split(x), y10); break;
      default: r10 = x10; break;
    }
    if (op < 17 || op == 25 || op == 26) r = fe10_canon(r10);
  }
  for (int l = 0; l < 8; ++l) out[l] = r.v[l];
}

// Ops 21 and 22: products by a wave-uniform constant as the search loop takes
// them from its table, the constants being the operands of the element that
// leads the element's group of 64 (i & ~63), A = its a and B = its b:
// op 21 is x * B, op 22 is x * B + y * A (one double product).
__host__ __device__ inline void field_op_const(int op, const uint32_t *a, const uint32_t *b,
                                               const uint32_t *A, const uint32_t *B,
                                               uint32_t *out) {
  Fe x, y, ca, cb;
  for (int l = 0; l < 8; ++l) {
    x.v[l] = a[l];
    y.v[l] = b[l];
    ca.v[l] = A[l];
    cb.v[l] = B[l];
  }
  const Fe10 x10 = fe10_from(x), y10 = fe10_from(y), a10 = fe10_from(ca), b10 = fe10_from(cb);
  const Fe r = fe10_canon(op == 21 ? fe10_mulc(x10, b10) : fe10_mulc2(x10, b10, y10, a10));
  for (int l = 0; l < 8; ++l) out[l] = r.v[l];
}

__global__ __launch_bounds__(256) void og_field_kernel(int op, const uint32_t *a, const uint32_t *b,
                                                       uint32_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 21 || op == 22) {
    // every wave's first lane is element i & ~63, and it is active
    uint32_t A[8], B[8];
    for (int l = 0; l < 8; ++l) {
      A[l] = __builtin_amdgcn_readfirstlane(a[(size_t)i * 8 + l]);
      B[l] = __builtin_amdgcn_readfirstlane(b[(size_t)i * 8 + l]);
    }
    field_op_const(op, a + (size_t)i * 8, b + (size_t)i * 8, A, B, out + (size_t)i * 8);
    return;
  }
  field_op_one(op, a + (size_t)i * 8, b + (size_t)i * 8, out + (size_t)i * 8);
}

// Radix bench: each lane runs two independent chains of `iters` dependent
// multiplications x <- x*y in the chosen representation (5: 5 x 51 in 64-bit
// limbs, 8: 8 x 32 saturated, 9: 9 x 29, 10: 10 x 25.5), then writes the
// canonical results.  Inputs are a fixed
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
  } else if (RADIX == 5) {
    Fe51 a5 = fe51_from(a), b5 = fe51_from(b), y5 = fe51_from(y);
    for (uint32_t i = 0; i < iters; ++i) {
      a5 = fe51_mul(a5, y5);
      b5 = fe51_mul(b5, y5);
    }
    a = fe51_to(a5);
    b = fe51_to(b5);
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

// Issue-rate probe: every lane runs eight independent chains of one VALU
// instruction, unrolled eight deep, so that the SIMD always has an
// instruction ready; the rate is lane-instructions per second.  Lane 0 of
// each workgroup's first wave also reads the shader clock (s_memtime) and the
// constant-rate clock (s_memrealtime) around the loop, from which the host
// derives the clock the SIMDs ran at.  OP 0: v_mad_u64_u32 (the product the
// field multiplication issues 100 of), 1: v_mul_lo_u32, 2: v_lshrrev_b64,
// 3: v_add_u32, 4: v_bitop3_b32 (a three-input XOR) and 5: v_alignbit_b32 (a
// 32-bit half of a 64-bit rotation), the two instructions the Keccak rounds
// are made of; 6: v_mad_u64_u32 in one dependent chain per lane, at three
// waves per SIMD (52 KiB of LDS per workgroup), the search loop's occupancy:
// whether a column of dependent products stalls the SIMD.
template <int OP>
__global__ __launch_bounds__(256) void og_issue_kernel(uint32_t iters, uint64_t *out,
                                                       uint64_t *clocks) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ uint32_t pad[OP == 6 ? 13312 : 1];
  if (OP == 6 && iters == 0xffffffffu) pad[threadIdx.x % (OP == 6 ? 13312 : 1)] = g;  // keeps it
  uint64_t acc[8];
  uint32_t a[8], b[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    a[c] = g * 0x9E3779B9u + (uint32_t)c * 0x85EBCA6Bu + 1u;
    b[c] = g ^ (0x27220A95u * (uint32_t)(c + 1));
    acc[c] = ((uint64_t)b[c] << 32) | a[c];
  }
  uint64_t t0 = 0, r0 = 0;
  if (threadIdx.x == 0) {
    t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (OP == 0) {
          uint64_t cc;
          asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0"
                       : "+v"(acc[c]), "=s"(cc)
                       : "v"(a[c]), "v"(b[c]));
        } else if (OP == 6) {
          uint64_t cc;
          asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0"
                       : "+v"(acc[0]), "=s"(cc)
                       : "v"(a[0]), "v"(b[0]));
        } else if (OP == 1) {
          uint32_t lo = (uint32_t)acc[c];
          asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(lo) : "v"(b[c]));
          acc[c] = (acc[c] & 0xffffffff00000000ull) | lo;
        } else if (OP == 2) {
          asm volatile("v_lshrrev_b64 %0, 1, %0" : "+v"(acc[c]));
        } else if (OP == 4) {
          uint32_t lo = (uint32_t)acc[c];
          asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(lo) : "v"(a[c]), "v"(b[c]));
          acc[c] = (acc[c] & 0xffffffff00000000ull) | lo;
        } else if (OP == 5) {
          uint32_t lo = (uint32_t)acc[c];
          asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(lo) : "v"(b[c]));
          acc[c] = (acc[c] & 0xffffffff00000000ull) | lo;
        } else {
          uint32_t lo = (uint32_t)acc[c];
          asm volatile("v_add_u32 %0, %0, %1" : "+v"(lo) : "v"(b[c]));
          acc[c] = (acc[c] & 0xffffffff00000000ull) | lo;
        }
      }
    }
  }
  if (threadIdx.x == 0) {
    const uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    clocks[2 * blockIdx.x] = t1 - t0;
    clocks[2 * blockIdx.x + 1] = r1 - r0;
  }
  uint64_t x = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) x ^= acc[c];
  if (OP == 6 && iters == 0xffffffffu) x ^= pad[(threadIdx.x + 1) % (OP == 6 ? 13312 : 1)];
  out[g] = x;
}

}  // namespace

// -------------------------------------------------------------------- host

struct og_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t threads = 0, half = 0, rounds = 1, block = 256, cap = 0;
  uint64_t launches = 0;
  uint32_t *d_state = nullptr, *d_table = nullptr, *d_scratch = nullptr;
  DevFilterHost filt;
  uint32_t *d_counter = nullptr;
  HitRec *d_hits = nullptr;
  uint32_t *d_parked = nullptr;  // per chain, for the address kernel
  uint32_t *h_counter = nullptr;
  HitRec *h_hits = nullptr;
  SearchArgs args;
  float last_ms = 0;
  uint64_t bytes = 0;  // device memory the engine holds
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
  e->filt.free_device();
  (void)hipFree(e->d_counter);
  (void)hipFree(e->d_hits);
  (void)hipFree(e->d_parked);
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
  if ((MODE & kModePattern) && e->args.parked) {
    const bool sign = e->args.f.defer_sign != 0;
    if (e->args.f.hash_addr) {
      const uint32_t block = e->threads % 1024 == 0 ? 1024 : e->block;
      const dim3 g(e->threads / block), b(block);
      if (sign)
        hipLaunchKernelGGL((og_address_kernel<true, true>), g, b, 0, e->stream, e->args);
      else
        hipLaunchKernelGGL((og_address_kernel<false, true>), g, b, 0, e->stream, e->args);
    } else if (sign) {
      hipLaunchKernelGGL((og_address_kernel<true, false>), dim3(grid), dim3(e->block), 0,
                         e->stream, e->args);
    } else {
      hipLaunchKernelGGL((og_address_kernel<false, false>), dim3(grid), dim3(e->block), 0,
                         e->stream, e->args);
    }
  }
  return hipGetLastError();
}

hipError_t launch_mode(og_engine *e) {
  switch (e->filt.mode) {
    case kModePrefix: return launch_search<kModePrefix>(e);
    case kModeBitmap: return launch_search<kModeBitmap>(e);
    case kModePattern: return launch_search<kModePattern>(e);
    case kModePrefix | kModePattern: return launch_search<kModePrefix | kModePattern>(e);
    case kModeBitmap | kModePattern: return launch_search<kModeBitmap | kModePattern>(e);
    default: return hipErrorInvalidValue;
  }
}

int check_device(int ordinal, char *err, size_t errcap) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    put_err(err, errcap, "no HIP device is visible");
    return -1;
  }
  if (ordinal < 0 || ordinal >= count) {
    put_err(err, errcap, "device ordinal out of range");
    return -1;
  }
  return 0;
}

// The chance that one candidate survives the device's filters: for a
// bitmap alone its occupancy, for a filter set the sum over its filters of the
// chance each matches (placements times the share of symbols each position
// allows).
double survivor_rate(const FilterSet *fs, const DevFilterHost &filt) {
  if (!fs) {
    if (!filt.bits || filt.bitmap.empty()) return 0;
    uint64_t set = 0;
    for (uint32_t w : filt.bitmap) set += (uint64_t)__builtin_popcount(w);
    return (double)set / (double)(1ull << filt.bits);
  }
  double rate = 0;
  for (const Filter &f : fs->filters) {
    if (f.regex) {
      // its match rate is not computable from the text (src/filter.rs:1586-1590);
      // the proxies' pass rate bounds it
      rate += f.regex->proxy_rate();
      continue;
    }
    // per placement, symbol by symbol: uniform, except that symbol 55 is
    // always 'd' and symbol 54 one of a, i, q, y (src/key.rs:309-327)
    const int len = (int)f.sets.size();
    const int first = f.anchor == kEnd ? kAddressLen - len : 0;
    const int last = f.anchor == kStart ? 0 : kAddressLen - len;
    for (int at = first; at <= last; ++at) {
      double p = 1;
      for (int i = 0; i < len; ++i) {
        const uint32_t set = f.sets[i];
        const int j = at + i;
        if (j == 55) p *= (set >> 3) & 1u;
        else if (j == 54) p *= __builtin_popcount(set & 0x01010101u) / 4.0;
        else p *= __builtin_popcount(set) / 32.0;
      }
      rate += p;
    }
  }
  return rate;
}

// Survivor room per launch: the reference's hit_room (src/gpu/mod.rs:306-312),
// raised to four times the expected survivors plus 1024 when the filters let
// more through than one bitmap slot in 2^bits, up to 2^22 records.
uint32_t survivor_room(uint64_t batch, const FilterSet *fs, const DevFilterHost &filt) {
  const uint32_t floor_room = og_hit_room(batch, filt.bits ? filt.bits : 24);
  const double want = 4.0 * (double)batch * survivor_rate(fs, filt) + 1024.0;
  return (uint32_t)std::max<double>(floor_room, std::min<double>(want, (double)(1u << 22)));
}

// Device bytes an engine of this shape holds: chain state (4 coordinates of
// 10 limbs), the offset table, the scratch planes (the pattern kernels' with
// their parking planes), the survivor ring, the parked counts of a one-round
// pattern launch and the filter tables.
uint64_t shape_bytes(uint32_t threads, uint32_t half, uint32_t rounds, uint32_t cap, bool pattern,
                     uint64_t filter_bytes) {
  return (uint64_t)160 * threads + (uint64_t)kTableWords * 4 * std::max<uint32_t>(half, 1) +
         (uint64_t)scratch_planes(half, pattern) * 40 * threads + 4 +
         (uint64_t)cap * sizeof(HitRec) + (pattern && rounds == 1 ? (uint64_t)4 * threads : 0) +
         filter_bytes;
}

// The memory an engine may take when its shape is chosen for it: the
// caller's cap, else half of what the device has free.
uint64_t engine_budget(int ordinal, uint64_t cap) {
  if (cap) return cap;
  size_t free_b = 0, total_b = 0;
  if (hipSetDevice(ordinal) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
    return 0;
  return (uint64_t)free_b / 2;
}

int create_engine(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                  const FilterSet *fs, const uint64_t *bitmap, size_t bitmap_words,
                  uint32_t bitmap_bits, og_engine **out, char *err, size_t errcap) {
  if (!out || !start_xyzt) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  *out = nullptr;
  if (check_device(ordinal, err, errcap)) return -1;
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
  if (fs) {
    std::string msg = e->filt.compile(*fs);
    if (!msg.empty()) {
      put_err(err, errcap, msg);
      return -1;
    }
  } else {
    if (!bitmap || bitmap_bits < 8 || bitmap_bits > 32 ||
        bitmap_words < (((size_t)1 << bitmap_bits) + 63) / 64) {
      put_err(err, errcap, "a bitmap of 8..32 bits with all its words is required");
      return -1;
    }
    e->filt.mode = kModeBitmap;
    e->filt.bits = bitmap_bits;
    e->filt.set_bitmap(std::vector<uint64_t>(bitmap, bitmap + (((size_t)1 << bitmap_bits) + 63) / 64));
  }
  OG_TRY(e->filt.to_device(e->stream));

  // ---- constants
  const uint64_t S = 2ull * o.half + 1;
  HostCurve hc;
  const std::vector<uint32_t> table = hc.offset_table10(o.half);
  const Cached1 stride = hc.multiple_of_eight_cached1((uint64_t)o.threads * S);
  SearchArgs &a = e->args;
  std::memset(&a, 0, sizeof a);
  put_fe10(a.stride, stride.ypx);
  put_fe10(a.stride + 10, stride.ymx);
  put_fe10(a.stride + 20, stride.t2d);

  // ---- buffers
  const size_t scratch_words =
      scratch_planes(o.half, (e->filt.mode & kModePattern) != 0) * 10 * o.threads;
  e->cap = o.hit_cap ? o.hit_cap
                     : survivor_room((uint64_t)o.threads * S * o.rounds, fs, e->filt);
  OG_TRY(hipMalloc(&e->d_state, (size_t)40 * o.threads * 4));
  OG_TRY(hipMalloc(&e->d_table, std::max<size_t>(table.size(), kTableWords) * 4));
  OG_TRY(hipMalloc(&e->d_scratch, std::max<size_t>(scratch_words, 1) * 4));
  OG_TRY(hipMalloc(&e->d_counter, 4));
  OG_TRY(hipMalloc(&e->d_hits, (size_t)e->cap * sizeof(HitRec)));
  OG_TRY(hipHostMalloc(&e->h_counter, 4, hipHostMallocDefault));
  OG_TRY(hipHostMalloc(&e->h_hits, (size_t)e->cap * sizeof(HitRec), hipHostMallocDefault));
  if (!table.empty())
    OG_TRY(hipMemcpyAsync(e->d_table, table.data(), table.size() * 4, hipMemcpyHostToDevice,
                          e->stream));
  a.state = e->d_state;
  a.table = e->d_table;
  a.scratch = e->d_scratch;
  a.threads = o.threads;
  a.half = o.half;
  a.rounds = o.rounds;
  a.f = e->filt.args();
  a.counter = e->d_counter;
  // A one-round pattern launch leaves its address stage to its own kernel,
  // and, unless many lanes park (the key stage run on sample keys), the sign
  // of x as well: then the search loop never forms x.
  if ((e->filt.mode & kModePattern) && o.rounds == 1) {
    OG_TRY(hipMalloc(&e->d_parked, (size_t)o.threads * 4));
    a.parked = e->d_parked;
    // the share that parks: the key stage's, plus the window prefilter's hits
    const double parks = e->filt.need + (e->filt.hash_key ? e->filt.window_park : 0.0);
    if (o.half >= 8 && o.sign_stage != 1)
      a.f.defer_sign = o.sign_stage == 2 || parks < 0.125 ? 1u : 0u;
  } else {
    a.parked = nullptr;
  }
  a.hits = e->d_hits;
  a.cap = e->cap;
  e->bytes = shape_bytes(o.threads, o.half, o.rounds, e->cap, (e->filt.mode & kModePattern) != 0,
                         e->filt.device_bytes());

  // ---- per-chain start points on the device
  SetupArgs sa = hc.setup_args(start_xyzt, e->d_state, o.threads, 0, o.threads, o.half);
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

}  // extern "C"

namespace {

// The cache key of this device's search shape: device/<name>/hip/0/<version>
// (src/tuning.rs:65-74, the threads field 0 as the reference's automatic
// choice records it).
std::string tuning_key(int ordinal) {
  char name[256] = {0};
  uint64_t mem = 0;
  int cus = 0;
  og_hip_device_info(ordinal, name, sizeof name, &mem, &cus);
  std::string key(og_tuning_device_key(name, "hip", 0, nullptr, 0), '\0');
  og_tuning_device_key(name, "hip", 0, &key[0], key.size() + 1);
  return key;
}

// The shapes the walk times, largest first (src/gpu/mod.rs:315-340 walks its
// batch sizes down the same way): chains in flight times pairs per window,
// with enough rounds for about 2^28 candidates per launch.  The largest,
// 2^20 chains of 1024 pairs, holds 128 GB of scratch planes.
std::vector<og_engine_opts> tuning_walk() {
  std::vector<og_engine_opts> out;
  for (uint32_t chains : {1u << 20, 1u << 19, 1u << 18, 1u << 17, 1u << 16})
    for (uint32_t half : {1024u, 512u, 256u, 128u}) {
      og_engine_opts o{};
      o.threads = chains;
      o.half = half;
      o.rounds = std::max<uint32_t>(1, (uint32_t)((1ull << 28) / ((uint64_t)chains * (2 * half + 1))));
      o.hit_cap = 1u << 20;
      out.push_back(o);
    }
  return out;
}

// The filter tables' device bytes for a shape estimate.
uint64_t filter_bytes(const FilterSet *fs, size_t bitmap_words) {
  if (!fs) return (uint64_t)bitmap_words * 8;
  DevFilterHost h;
  return h.compile(*fs).empty() ? h.device_bytes() : 0;
}

// Times one launch of every shape that fits `budget` bytes after a warm-up
// launch (which pays for the tables reaching the device), as
// src/gpu/mod.rs:415-428 does, in candidates per second of wall time.  The
// reference bounds its walk by memory too (src/gpu/mod.rs:314-342): here
// shapes above the budget are not built, and among the shapes within 3% of
// the fastest the smallest footprint is kept (og_tuning_choose), so that an
// engine never holds most of the device for a rate it would nearly reach with
// a fraction of it.  The choice is recorded.
int tune(int ordinal, const uint8_t start_xyzt[128], const FilterSet *fs, const uint64_t *bitmap,
         size_t bitmap_words, uint32_t bitmap_bits, uint64_t budget, og_tuning_shape *best,
         char *err, size_t errcap) {
  og_tuning_shape b{};
  std::string last = "no shape was tried";
  const std::vector<og_engine_opts> walk = tuning_walk();
  std::vector<double> rates(walk.size(), 0.0);
  std::vector<uint64_t> sizes(walk.size(), 0);
  const uint64_t fbytes = filter_bytes(fs, bitmap_words);
  const bool pattern = fs && (!fs->general.empty() || !fs->regexes.empty());
  for (size_t i = 0; i < walk.size(); ++i) {
    const og_engine_opts &o = walk[i];
    sizes[i] = shape_bytes(o.threads, o.half, o.rounds, o.hit_cap, pattern, fbytes);
    if (budget && sizes[i] > budget) {
      last = "every shape exceeds the memory budget of " + std::to_string(budget) + " bytes";
      continue;
    }
    og_engine *e = nullptr;
    char e2[512] = {0};
    if (create_engine(ordinal, start_xyzt, &o, fs, bitmap, bitmap_words, bitmap_bits, &e, e2,
                      sizeof e2) != 0) {
      last = e2;
      continue;
    }
    std::vector<og_survivor> room(e->cap);
    uint32_t n = 0;
    double rate = 0;
    if (og_hip_engine_advance(e, room.data(), e->cap, &n, e2, sizeof e2) == 0) {
      const auto t0 = std::chrono::steady_clock::now();
      if (og_hip_engine_advance(e, room.data(), e->cap, &n, e2, sizeof e2) == 0) {
        const double sec =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (sec > 0) rate = (double)og_hip_engine_batch(e) / sec;
      }
    }
    if (rate == 0 && e2[0]) last = e2;
    if (rate > 0) sizes[i] = e->bytes;
    og_hip_engine_destroy(e);
    rates[i] = rate;
  }
  const size_t pick = og_tuning_choose(rates.data(), sizes.data(), walk.size(), budget, 0.03);
  if (pick < walk.size()) {
    b.first = walk[pick].half;
    b.second = walk[pick].threads;
    b.rounds = walk[pick].rounds;
    b.rate = rates[pick];
  }
  if (b.rate == 0) {
    put_err(err, errcap, "no search shape this device would take: " + last);
    return -1;
  }
  b.at = (int64_t)std::time(nullptr);
  b.radix = 10;
  og_tuning_put(tuning_key(ordinal).c_str(), &b);
  if (best) *best = b;
  return 0;
}

// threads == 0: the whole shape from the tuning cache, else measured now
// (src/gpu/mod.rs:354-412).  A cached shape that no longer builds is measured
// again.
int create_tuned(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                 const FilterSet *fs, const uint64_t *bitmap, size_t bitmap_words,
                 uint32_t bitmap_bits, og_engine **out, char *err, size_t errcap) {
  if (opts && opts->threads)
    return create_engine(ordinal, start_xyzt, opts, fs, bitmap, bitmap_words, bitmap_bits, out,
                         err, errcap);
  if (check_device(ordinal, err, errcap)) return -1;
  og_engine_opts o{};
  if (opts) {
    o.hit_cap = opts->hit_cap;
    o.sign_stage = opts->sign_stage;
  }
  const uint64_t budget = engine_budget(ordinal, opts ? opts->max_bytes : 0);
  og_tuning_shape s{};
  if (og_tuning_get(tuning_key(ordinal).c_str(), &s) == 1 && s.first && s.second) {
    o.threads = s.second;
    o.half = s.first;
    o.rounds = s.rounds ? s.rounds : 1;
    // a remembered shape above this engine's budget is measured again, as one
    // that no longer builds is (src/gpu/mod.rs:371-381)
    const bool pattern = fs && (!fs->general.empty() || !fs->regexes.empty());
    const uint64_t est = shape_bytes(o.threads, o.half, o.rounds,
                                     o.hit_cap ? o.hit_cap : (1u << 20), pattern,
                                     filter_bytes(fs, bitmap_words));
    if ((!budget || est <= budget) &&
        create_engine(ordinal, start_xyzt, &o, fs, bitmap, bitmap_words, bitmap_bits, out, err,
                      errcap) == 0)
      return 0;
  }
  if (tune(ordinal, start_xyzt, fs, bitmap, bitmap_words, bitmap_bits, budget, &s, err,
           errcap) != 0)
    return -1;
  o.threads = s.second;
  o.half = s.first;
  o.rounds = s.rounds;
  return create_engine(ordinal, start_xyzt, &o, fs, bitmap, bitmap_words, bitmap_bits, out, err,
                       errcap);
}

}  // namespace

extern "C" {

int og_hip_engine_create(int ordinal, const uint8_t start_xyzt[128], uint32_t threads,
                         const uint64_t *bitmap, size_t bitmap_words, uint32_t bitmap_bits,
                         og_engine **out, char *err, size_t errcap) {
  og_engine_opts o{};
  o.threads = threads;
  return create_tuned(ordinal, start_xyzt, &o, nullptr, bitmap, bitmap_words, bitmap_bits, out,
                      err, errcap);
}

int og_hip_engine_create_ex(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                            const og_filterset *filters, og_engine **out, char *err,
                            size_t errcap) {
  if (!filters) {
    put_err(err, errcap, "null filter set");
    return -1;
  }
  return create_tuned(ordinal, start_xyzt, opts, filterset_of(filters), nullptr, 0, 0, out, err,
                      errcap);
}

int og_hip_engine_tune(int ordinal, const uint8_t start_xyzt[128], const og_filterset *filters,
                       uint64_t max_bytes, og_tuning_shape *best, char *err, size_t errcap) {
  if (!start_xyzt || !filters) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  if (check_device(ordinal, err, errcap)) return -1;
  return tune(ordinal, start_xyzt, filterset_of(filters), nullptr, 0, 0,
              engine_budget(ordinal, max_bytes), best, err, errcap);
}

uint64_t og_hip_engine_bytes(const og_engine *e) { return e ? e->bytes : 0; }

void og_hip_engine_shape(const og_engine *e, uint32_t *threads, uint32_t *half,
                         uint32_t *rounds) {
  if (threads) *threads = e ? e->threads : 0;
  if (half) *half = e ? e->half : 0;
  if (rounds) *rounds = e ? e->rounds : 0;
}

int og_hip_engine_sign_stage(const og_engine *e) {
  if (!e) return -1;
  return e->args.parked && e->args.f.defer_sign ? 2 : 1;
}

int og_hip_engine_kernels(const og_engine *e) {
  if (!e) return -1;
  return (e->filt.mode & kModePattern) ? 2 : 1;
}

uint32_t og_hip_engine_threads(const og_engine *e) { return e ? e->threads : 0; }

uint32_t og_hip_engine_room(const og_engine *e) { return e ? e->cap : 0; }

int og_hip_engine_device(const og_engine *e) { return e ? e->device : -1; }

int og_hip_engine_restart(og_engine *e, const uint8_t start_xyzt[128], char *err, size_t errcap) {
  if (!e || !start_xyzt) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  OG_TRY(hipSetDevice(e->device));
  HostCurve hc;
  SetupArgs sa = hc.setup_args(start_xyzt, e->d_state, e->threads, 0, e->threads, e->half);
  hipLaunchKernelGGL(og_setup_kernel, dim3(e->threads / 256), dim3(256), 0, e->stream, sa);
  OG_TRY(hipGetLastError());
  OG_TRY(hipStreamSynchronize(e->stream));
  e->launches = 0;
  return 0;
}

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
  if ((radix != 5 && radix != 8 && radix != 9 && radix != 10) || threads == 0 || threads % 256 ||
      threads > (1u << 22)) {
    put_err(err, errcap, "radix must be 5, 8, 9 or 10 and threads a multiple of 256 up to 2^22");
    return -1;
  }
  OG_TRY(hipSetDevice(ordinal));
  uint32_t *d = nullptr;
  OG_TRY(hipMalloc(&d, (size_t)threads * 64));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&](uint32_t it) {
    if (radix == 5)
      hipLaunchKernelGGL(og_radix_kernel<5>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
    else if (radix == 8)
      hipLaunchKernelGGL(og_radix_kernel<8>, dim3(threads / 256), dim3(256), 0, 0, it, d, threads);
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

int og_hip_issue_bench(int ordinal, int op, uint32_t threads, uint32_t iters, double *ops_per_s,
                       double *clock_hz, char *err, size_t errcap) {
  if (op < 0 || op > 6 || threads == 0 || threads % 256 || threads > (1u << 22) || iters == 0) {
    put_err(err, errcap, "op must be 0..6, threads a multiple of 256 up to 2^22, iters positive");
    return -1;
  }
  OG_TRY(hipSetDevice(ordinal));
  int wall_khz = 0;
  OG_TRY(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, ordinal));
  const uint32_t blocks = threads / 256;
  uint64_t *d = nullptr, *clk = nullptr;
  OG_TRY(hipMalloc(&d, (size_t)threads * 8));
  hipError_t e = hipMalloc(&clk, (size_t)blocks * 16);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  auto launch = [&](uint32_t it) {
    switch (op) {
      case 0: hipLaunchKernelGGL(og_issue_kernel<0>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      case 1: hipLaunchKernelGGL(og_issue_kernel<1>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      case 2: hipLaunchKernelGGL(og_issue_kernel<2>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      case 4: hipLaunchKernelGGL(og_issue_kernel<4>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      case 5: hipLaunchKernelGGL(og_issue_kernel<5>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      case 6: hipLaunchKernelGGL(og_issue_kernel<6>, dim3(blocks), dim3(256), 0, 0, it, d, clk); break;
      default: hipLaunchKernelGGL(og_issue_kernel<3>, dim3(blocks), dim3(256), 0, 0, it, d, clk);
    }
  };
  float ms = 0;
  std::vector<uint64_t> h(2 * (size_t)blocks);
  if (e == hipSuccess) {
    launch(iters < 64 ? iters : 64);  // warm-up
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  if (e == hipSuccess) launch(iters);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) e = hipMemcpy(h.data(), clk, h.size() * 8, hipMemcpyDeviceToHost);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(d);
  if (clk) (void)hipFree(clk);
  if (e != hipSuccess) {
    put_err(err, errcap, std::string("issue bench: ") + hipGetErrorString(e));
    return -1;
  }
  if (ops_per_s) *ops_per_s = 64.0 * threads * (double)iters / (ms * 1e-3);
  if (clock_hz) {
    double ticks = 0, real = 0;
    for (uint32_t i = 0; i < blocks; ++i) {
      ticks += (double)h[2 * i];
      real += (double)h[2 * i + 1];
    }
    *clock_hz = real > 0 && wall_khz > 0 ? ticks / real * wall_khz * 1e3 : 0.0;
  }
  return 0;
}

// ops 0..5, 10..26, as include/onion_gen_hip.h lists them
static bool field_op_known(int op) { return (op >= 0 && op <= 5) || (op >= 10 && op <= 27); }

int og_field_op_host(int op, const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
  if (!field_op_known(op) || !a || !b || !out) return -1;
  for (size_t i = 0; i < n; ++i) {
    uint32_t x[8], y[8], r[8];
    std::memcpy(x, a + 32 * i, 32);
    std::memcpy(y, b + 32 * i, 32);
    if (op == 21 || op == 22) {
      uint32_t A[8], B[8];
      std::memcpy(A, a + 32 * (i & ~(size_t)63), 32);
      std::memcpy(B, b + 32 * (i & ~(size_t)63), 32);
      field_op_const(op, x, y, A, B, r);
    } else {
      field_op_one(op, x, y, r);
    }
    std::memcpy(out + 32 * i, r, 32);
  }
  return 0;
}

int og_hip_field_op(int ordinal, int op, const uint8_t *a, const uint8_t *b, uint8_t *out,
                    size_t n, char *err, size_t errcap) {
  if (!field_op_known(op) || !a || !b || !out) {
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
