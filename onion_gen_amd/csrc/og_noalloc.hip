// One search launch that allocates nothing: for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops that must not
// allocate, harnesses that forbid it).  It runs what one advance of an engine
// of the same shape runs (og_hip_engine_create_ex + og_hip_engine_advance,
// one window per chain) for one literal prefix, from a given candidate index,
// with every buffer the library would otherwise allocate - the offset table,
// the chain state, the scratch planes, the survivor ring and its counter - in
// one static device array of kNoallocBytes per device.  The caller passes the
// inputs and device memory for the outputs; the work is ordered on the
// caller's stream and nothing waits for it.
//
// Every call on a device shares that device's array, and the array keeps the
// last call's contents: each call clears the counter first, and calls that
// overlap on one device race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/onion_gen_hip.h"
#include "og_devfilter.hpp"
#include "og_filters.hpp"
#include "og_host.hpp"
#include "og_search.hpp"

using namespace og;

namespace {
constexpr size_t kNoallocBytes = (size_t)64 << 20;
constexpr uint32_t kNoallocRing = 16384;  // survivor records the array holds
constexpr int kMaxDevices = 64;
}  // namespace

// The workspace: each device holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

struct TableArgs {
  uint32_t eight[32];  // 8G cached: y+x, y-x, Z, 2dT (8 x 32)
  uint32_t d[8];
  uint32_t *table;     // [half][kTableWords]: x, y, d x y and the derived terms
  uint32_t half;
  uint32_t *counter;
};

// The offset table the search reads, Q_m = (m + 1) * 8G as affine x, y and
// d x y in the 10-limb form, each entry on its own lane (the host engine
// builds the same canonical values on the CPU); lane 0 also clears the
// survivor counter.
__global__ __launch_bounds__(256) void og_noalloc_table_kernel(TableArgs a) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0) *a.counter = 0;
  if (m >= a.half) return;
  Cached eight;
  Fe d;
  for (int l = 0; l < 8; ++l) {
    eight.ypx.v[l] = a.eight[l];
    eight.ymx.v[l] = a.eight[8 + l];
    eight.Z.v[l] = a.eight[16 + l];
    eight.t2d.v[l] = a.eight[24 + l];
    d.v[l] = a.d[l];
  }
  const Point q = pt_mul_u64((uint64_t)m + 1, eight);
  const Fe zinv = fe_invert(q.Z);
  const Fe x = fe_canon(fe_mul(q.X, zinv)), y = fe_canon(fe_mul(q.Y, zinv));
  const Fe dxy = fe_canon(fe_mul(d, fe_mul(x, y)));
  Fe c[kTableTerms] = {x, y, dxy};
  offset_terms(c);
  for (uint32_t i = 0; i < kTableTerms; ++i) {
    const Fe10 r = fe10_split(c[i]);
    for (int l = 0; l < 10; ++l) a.table[(size_t)m * kTableWords + 10 * i + l] = r.v[l];
  }
}

__global__ __launch_bounds__(256) void og_noalloc_setup_kernel(SetupArgs a) {
  const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.count) return;
  setup_chain(a, local);
}

// One window per chain, the literal-prefix form: the search loop of the
// engine's kernel on this call's buffers.
__global__ __launch_bounds__(256) OG_SEARCH_ATTR(kModePrefix) void og_noalloc_search_kernel(
    SearchArgs a) {
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
  __shared__ ScanBank banks[1];
  __shared__ uint16_t dfa_lds[2];
  HitSink sink{a.counter, a.hits, a.cap, 0, nullptr, 0};
  search_chain<kModePrefix>(P, st, (cptr)a.table, a.scratch, T, g, a.half, 1, T, g, a.k_base,
                            a.f, banks, dfa_lds, nullptr, sink);
}

// The ring's records as the caller's og_survivor (offset = 8k, packed = the
// canonical y bytes), at most `cap` of them, and the count found.
__global__ __launch_bounds__(256) void og_noalloc_out_kernel(const uint32_t *counter,
                                                             const HitRec *hits, uint32_t ring,
                                                             og_survivor *out, uint32_t cap,
                                                             uint32_t *n_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t found = *counter;
  const uint32_t kept = found < cap ? (found < ring ? found : ring) : (cap < ring ? cap : ring);
  if (i == 0) *n_out = found;
  if (i >= kept) return;
  out[i].offset = hits[i].k * 8;
  uint32_t *p = reinterpret_cast<uint32_t *>(out[i].packed);
  for (int l = 0; l < 8; ++l) p[l] = hits[i].y[l];
}

void put_err(char *err, size_t cap, const std::string &msg) {
  if (err && cap) {
    const size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The device address of this device's array, looked up once per device with
// that device current.  The array's name in host code is a stand-in, never a
// device address: every use goes through the pointer kept here.
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices];

unsigned char *device_array(int ordinal, std::string &msg) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[ordinal]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      msg = std::string("hipGetSymbolAddress of the static workspace: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_ptr[ordinal] = static_cast<unsigned char *>(p);
  }
  return g_ptr[ordinal];
}

const HostCurve &curve() {
  static const HostCurve hc;
  return hc;
}

}  // namespace

extern "C" {

size_t og_hip_search_noalloc_bytes(uint32_t threads, uint32_t half) {
  return 256 + align256((size_t)kNoallocRing * sizeof(HitRec)) +
         align256((size_t)half * kTableWords * 4) + align256((size_t)160 * threads) +
         scratch_planes(half, false) * 40 * (size_t)threads;
}

size_t og_hip_search_noalloc_max(void) { return kNoallocBytes; }

int og_hip_search_noalloc(int ordinal, const uint8_t start_xyzt[128], uint64_t k0,
                          uint32_t threads, uint32_t half, const og_filterset *filters,
                          og_survivor *out, uint32_t cap, uint32_t *n_out, void *stream,
                          char *err, size_t errcap) {
  if (!start_xyzt || !filters || !n_out || (cap && !out)) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  if (ordinal < 0 || ordinal >= kMaxDevices) {
    put_err(err, errcap, "no such device");
    return -1;
  }
  if (threads == 0 || threads % 256 || threads > (1u << 22) || half == 0 || half > 4096) {
    put_err(err, errcap, "threads must be a positive multiple of 256 and half in 1..4096");
    return -1;
  }
  const uint64_t batch = (uint64_t)threads * (2ull * half + 1);
  if (k0 > UINT64_MAX / 8 - batch) {
    put_err(err, errcap, "candidate indices past 2^61 have no offset");
    return -1;
  }
  const size_t need = og_hip_search_noalloc_bytes(threads, half);
  if (need > kNoallocBytes) {
    put_err(err, errcap,
            "this shape needs " + std::to_string(need) + " bytes of workspace and the static one "
            "holds " + std::to_string(kNoallocBytes) + ": use fewer chains or pairs per window");
    return -1;
  }
  const FilterSet *fs = filterset_of(filters);
  if (!fs || fs->fast.size() != 1 || !fs->general.empty() || !fs->regexes.empty()) {
    put_err(err, errcap, "the allocation-free search takes exactly one literal prefix");
    return -1;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || ordinal >= count) {
    put_err(err, errcap, "no such device");
    return -1;
  }
  if (hipSetDevice(ordinal) != hipSuccess) {
    put_err(err, errcap, "hipSetDevice failed");
    return -1;
  }
  std::string msg;
  unsigned char *base = device_array(ordinal, msg);
  if (!base) {
    put_err(err, errcap, msg);
    return -1;
  }
  // the array's layout for this shape
  uint32_t *counter = reinterpret_cast<uint32_t *>(base);
  HitRec *ring = reinterpret_cast<HitRec *>(base + 256);
  size_t off = 256 + align256((size_t)kNoallocRing * sizeof(HitRec));
  uint32_t *table = reinterpret_cast<uint32_t *>(base + off);
  off += align256((size_t)half * kTableWords * 4);
  uint32_t *state = reinterpret_cast<uint32_t *>(base + off);
  off += align256((size_t)160 * threads);
  uint32_t *planes = reinterpret_cast<uint32_t *>(base + off);

  const HostCurve &hc = curve();
  const hipStream_t s = static_cast<hipStream_t>(stream);
  TableArgs ta;
  std::memset(&ta, 0, sizeof ta);
  const Cached eight = hc.eight_cached();
  for (int l = 0; l < 8; ++l) {
    ta.eight[l] = eight.ypx.v[l];
    ta.eight[8 + l] = eight.ymx.v[l];
    ta.eight[16 + l] = eight.Z.v[l];
    ta.eight[24 + l] = eight.t2d.v[l];
    ta.d[l] = hc.d.v[l];
  }
  ta.table = table;
  ta.half = half;
  ta.counter = counter;
  hipLaunchKernelGGL(og_noalloc_table_kernel, dim3((half + 255) / 256), dim3(256), 0, s, ta);

  const SetupArgs sa = hc.setup_args(start_xyzt, state, threads, 0, threads, half, k0);
  hipLaunchKernelGGL(og_noalloc_setup_kernel, dim3(threads / 256), dim3(256), 0, s, sa);

  SearchArgs a;
  std::memset(&a, 0, sizeof a);
  const Cached1 stride = hc.multiple_of_eight_cached1(batch);
  put_fe10(a.stride, stride.ypx);
  put_fe10(a.stride + 10, stride.ymx);
  put_fe10(a.stride + 20, stride.t2d);
  a.state = state;
  a.table = table;
  a.scratch = planes;
  a.threads = threads;
  a.half = half;
  a.rounds = 1;
  a.k_base = k0;
  a.f.mode = kModePrefix;
  prefix_words(fs->filters[fs->fast[0]], a.f.want, a.f.mask);
  a.counter = counter;
  a.hits = ring;
  a.cap = kNoallocRing;
  a.parked = nullptr;
  hipLaunchKernelGGL(og_noalloc_search_kernel, dim3(threads / 256), dim3(256), 0, s, a);

  const uint32_t writes = std::max<uint32_t>(1, std::min(cap, kNoallocRing));
  hipLaunchKernelGGL(og_noalloc_out_kernel, dim3((writes + 255) / 256), dim3(256), 0, s, counter,
                     ring, kNoallocRing, out, cap, n_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    put_err(err, errcap, std::string("launch: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

}  // extern "C"
