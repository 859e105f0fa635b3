// The slot engine: many independent searches in one launch.
//
// A batch holds `slots` search slots of `chains` chains each.  One step is a
// captured HIP graph - clear the survivor counter, run every chain of every
// active slot for `rounds` windows, advance each active slot's candidate base,
// copy the counter back - replayed as is, so admitting or retiring a search
// never re-captures anything: a slot's descriptor in device memory says
// whether its blocks work and which filter they apply.  Searches of different
// filter kinds share the launch; each block reads its slot's descriptor once.
//
// This is the batching the reference does not have (its device path runs one
// search per engine, src/gpu/mod.rs:292-477); per slot the arithmetic, the
// candidate numbering and the survivor contract are the single engine's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"
#include "og_devfilter.hpp"
#include "og_host.hpp"
#include "og_search.hpp"

using namespace og;

namespace {

struct SlotDesc {
  DevFilterArgs f;      // f.mode == 0: inactive
  uint64_t k_base;      // next candidate of the slot's search
};

struct BatchArgs {
  uint32_t *state;
  const uint32_t *table;
  uint32_t *scratch;
  SlotDesc *slots;
  uint32_t nslots, chains, half, rounds;
  uint32_t stride[30];
  uint32_t *counter;
  HitRec *hits;
  uint32_t cap;
  uint32_t *slot_hits;
};

template <uint32_t MODE>
__device__ __forceinline__ void batch_body(const BatchArgs &a, const SlotDesc &d, uint32_t slot,
                                           uint32_t local, uint32_t g, uint32_t T) {
  Point10 P = ld_point(a.state, T, g);
  Cached10 st;
#pragma unroll
  for (int l = 0; l < 10; ++l) {
    st.ypx.v[l] = a.stride[l];
    st.ymx.v[l] = a.stride[10 + l];
    st.t2d.v[l] = a.stride[20 + l];
  }
  HitSink sink{a.counter, a.hits, a.cap, slot, a.slot_hits};
  search_chain<MODE>(P, st, (cptr)a.table, a.scratch, T, g, a.half, a.rounds, a.chains, local,
                     d.k_base, d.f, sink);
  st_point(a.state, T, g, P);
}

__global__ __launch_bounds__(256) OG_SEARCH_ATTR void og_batch_kernel(BatchArgs a) {
  const uint32_t blocks_per_slot = a.chains / 256;
  const uint32_t slot = blockIdx.x / blocks_per_slot;
  const uint32_t local = (blockIdx.x % blocks_per_slot) * 256 + threadIdx.x;
  const uint32_t T = a.nslots * a.chains;
  const uint32_t g = slot * a.chains + local;
  const SlotDesc d = a.slots[slot];  // block-uniform
  switch (d.f.mode) {
    case kModePrefix: batch_body<kModePrefix>(a, d, slot, local, g, T); break;
    case kModeBitmap: batch_body<kModeBitmap>(a, d, slot, local, g, T); break;
    case kModePattern: batch_body<kModePattern>(a, d, slot, local, g, T); break;
    case kModePrefix | kModePattern: batch_body<kModePrefix | kModePattern>(a, d, slot, local, g, T); break;
    case kModeBitmap | kModePattern: batch_body<kModeBitmap | kModePattern>(a, d, slot, local, g, T); break;
    default: break;  // inactive slot
  }
}

// Clears the survivor counter (a kernel rather than a memset node: the graph
// is replayed many times).
__global__ void og_batch_clear(uint32_t *counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

// Moves every active slot on by one step's candidates.
__global__ void og_batch_advance(SlotDesc *slots, uint32_t nslots, uint64_t per_step) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nslots && slots[s].f.mode) slots[s].k_base += per_step;
}

__global__ __launch_bounds__(256) void og_batch_setup(SetupArgs a) {
  const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= a.count) return;
  setup_chain(a, local);
}

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

struct HostSlot {
  bool active = false;
  uint64_t tag = 0;
  uint64_t examined = 0;
  uint64_t hits = 0;
  DevFilterHost filt;
};

}  // namespace

struct og_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  og_batch_opts o{};
  uint64_t per_step = 0;  // candidates one step examines per active slot
  uint32_t *d_state = nullptr, *d_table = nullptr, *d_scratch = nullptr;
  SlotDesc *d_slots = nullptr;
  uint32_t *d_counter = nullptr, *d_slot_hits = nullptr;
  HitRec *d_hits = nullptr;
  uint32_t *h_counter = nullptr;
  HitRec *h_hits = nullptr;
  std::vector<HostSlot> slots;
  std::vector<uint8_t> start_cache;  // not used by the device; kept for diagnostics
  BatchArgs args;
  HostCurve curve;
  float last_ms = 0;
  uint64_t steps = 0;
  size_t table_bytes_budget = 0, table_bytes_used = 0;
};

namespace {

void free_batch(og_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (HostSlot &s : b->slots) s.filt.free_device();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->d_state);
  (void)hipFree(b->d_table);
  (void)hipFree(b->d_scratch);
  (void)hipFree(b->d_slots);
  (void)hipFree(b->d_counter);
  (void)hipFree(b->d_slot_hits);
  (void)hipFree(b->d_hits);
  (void)hipHostFree(b->h_counter);
  (void)hipHostFree(b->h_hits);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

extern "C" {

int og_batch_create(int ordinal, const og_batch_opts *opts, og_batch **out, char *err,
                    size_t errcap) {
  if (!out) {
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
  og_batch_opts o = opts ? *opts : og_batch_opts{};
  if (!o.slots) o.slots = 64;
  if (!o.chains) o.chains = 2048;
  if (!o.half) o.half = 256;
  if (!o.rounds) o.rounds = 1;
  if (!o.ring_cap) o.ring_cap = 1u << 16;
  if (o.chains % 256 || (uint64_t)o.slots * o.chains > (1u << 22) || o.half > 4096 ||
      o.slots > 4096) {
    put_err(err, errcap,
            "chains must be a multiple of 256, slots*chains at most 2^22, half at most 4096");
    return -1;
  }
  std::unique_ptr<og_batch, void (*)(og_batch *)> b(new og_batch(), free_batch);
  b->device = ordinal;
  b->o = o;
  b->slots.resize(o.slots);
  const uint64_t S = 2ull * o.half + 1;
  b->per_step = (uint64_t)o.chains * S * o.rounds;
  const size_t T = (size_t)o.slots * o.chains;
  OG_TRY(hipSetDevice(ordinal));
  OG_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  OG_TRY(hipEventCreate(&b->ev0));
  OG_TRY(hipEventCreate(&b->ev1));
  const std::vector<uint32_t> table = b->curve.offset_table10(o.half);
  OG_TRY(hipMalloc(&b->d_state, T * 40 * 4));
  OG_TRY(hipMalloc(&b->d_table, std::max<size_t>(table.size(), 30) * 4));
  OG_TRY(hipMalloc(&b->d_scratch, (size_t)2 * o.half * 10 * T * 2 * 4));
  OG_TRY(hipMalloc(&b->d_slots, o.slots * sizeof(SlotDesc)));
  OG_TRY(hipMalloc(&b->d_counter, 4));
  OG_TRY(hipMalloc(&b->d_slot_hits, o.slots * 4));
  OG_TRY(hipMalloc(&b->d_hits, (size_t)o.ring_cap * sizeof(HitRec)));
  OG_TRY(hipHostMalloc(&b->h_counter, 4, hipHostMallocDefault));
  OG_TRY(hipHostMalloc(&b->h_hits, (size_t)o.ring_cap * sizeof(HitRec), hipHostMallocDefault));
  if (!table.empty())
    OG_TRY(hipMemcpyAsync(b->d_table, table.data(), table.size() * 4, hipMemcpyHostToDevice,
                          b->stream));
  std::vector<SlotDesc> zero(o.slots);
  std::memset(zero.data(), 0, zero.size() * sizeof(SlotDesc));
  OG_TRY(hipMemcpyAsync(b->d_slots, zero.data(), zero.size() * sizeof(SlotDesc),
                        hipMemcpyHostToDevice, b->stream));
  OG_TRY(hipStreamSynchronize(b->stream));

  BatchArgs &a = b->args;
  std::memset(&a, 0, sizeof a);
  a.state = b->d_state;
  a.table = b->d_table;
  a.scratch = b->d_scratch;
  a.slots = b->d_slots;
  a.nslots = o.slots;
  a.chains = o.chains;
  a.half = o.half;
  a.rounds = o.rounds;
  const Cached1 stride = b->curve.multiple_of_eight_cached1((uint64_t)o.chains * S);
  put_fe10(a.stride, stride.ypx);
  put_fe10(a.stride + 10, stride.ymx);
  put_fe10(a.stride + 20, stride.t2d);
  a.counter = b->d_counter;
  a.hits = b->d_hits;
  a.cap = o.ring_cap;
  a.slot_hits = b->d_slot_hits;

  // Capture one step.
  OG_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(og_batch_clear, dim3(1), dim3(64), 0, b->stream, b->d_counter);
  hipLaunchKernelGGL(og_batch_kernel, dim3((unsigned)(T / 256)), dim3(256), 0, b->stream, a);
  hipLaunchKernelGGL(og_batch_advance, dim3((o.slots + 255) / 256), dim3(256), 0, b->stream,
                     b->d_slots, o.slots, b->per_step);
  (void)hipMemcpyAsync(b->h_counter, b->d_counter, 4, hipMemcpyDeviceToHost, b->stream);
  OG_TRY(hipStreamEndCapture(b->stream, &b->graph));
  OG_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  *out = b.release();
  return 0;
}

size_t og_batch_footprint(const og_filterset *filters) {
  const FilterSet *fs = filterset_of(filters);
  if (!fs) return 0;
  DevFilterHost h;
  if (!h.compile(*fs).empty()) return 0;
  return h.device_bytes();
}

int og_batch_admit(og_batch *b, const uint8_t start_xyzt[128], const og_filterset *filters,
                   uint64_t tag, int *slot_out, char *err, size_t errcap) {
  if (!b || !start_xyzt || !filters || !slot_out) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  int slot = -1;
  for (size_t i = 0; i < b->slots.size(); ++i)
    if (!b->slots[i].active) {
      slot = (int)i;
      break;
    }
  if (slot < 0) {
    put_err(err, errcap, "no free slot");
    return -2;
  }
  HostSlot &hs = b->slots[slot];
  hs = HostSlot();
  std::string msg = hs.filt.compile(*filterset_of(filters));
  if (!msg.empty()) {
    put_err(err, errcap, msg);
    return -1;
  }
  OG_TRY(hipSetDevice(b->device));
  OG_TRY(hs.filt.to_device(b->stream));
  SlotDesc d;
  std::memset(&d, 0, sizeof d);
  d.f = hs.filt.args();
  d.k_base = 0;
  SetupArgs sa = b->curve.setup_args(start_xyzt, b->d_state, b->o.slots * b->o.chains,
                                     (uint32_t)slot * b->o.chains, b->o.chains, b->o.half);
  hipLaunchKernelGGL(og_batch_setup, dim3(b->o.chains / 256), dim3(256), 0, b->stream, sa);
  OG_TRY(hipGetLastError());
  uint32_t zero = 0;
  OG_TRY(hipMemcpyAsync(b->d_slot_hits + slot, &zero, 4, hipMemcpyHostToDevice, b->stream));
  OG_TRY(hipMemcpyAsync(b->d_slots + slot, &d, sizeof d, hipMemcpyHostToDevice, b->stream));
  OG_TRY(hipStreamSynchronize(b->stream));
  hs.active = true;
  hs.tag = tag;
  *slot_out = slot;
  return 0;
}

int og_batch_retire(og_batch *b, int slot) {
  if (!b || slot < 0 || (size_t)slot >= b->slots.size()) return -1;
  HostSlot &hs = b->slots[slot];
  if (!hs.active) return 0;
  (void)hipSetDevice(b->device);
  SlotDesc d;
  std::memset(&d, 0, sizeof d);
  if (hipMemcpyAsync(b->d_slots + slot, &d, sizeof d, hipMemcpyHostToDevice, b->stream) !=
      hipSuccess)
    return -1;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
  hs.filt.free_device();
  hs.active = false;
  return 0;
}

int og_batch_step(og_batch *b, og_batch_hit *out, uint32_t cap, uint32_t *n, char *err,
                  size_t errcap) {
  if (!b || !n) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  *n = 0;
  OG_TRY(hipSetDevice(b->device));
  OG_TRY(hipEventRecord(b->ev0, b->stream));
  OG_TRY(hipGraphLaunch(b->exec, b->stream));
  OG_TRY(hipEventRecord(b->ev1, b->stream));
  OG_TRY(hipStreamSynchronize(b->stream));
  OG_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  b->steps++;
  for (HostSlot &s : b->slots)
    if (s.active) s.examined += b->per_step;
  const uint32_t found = *b->h_counter;
  const uint32_t kept = std::min(found, b->o.ring_cap);
  if (kept) {
    OG_TRY(hipMemcpyAsync(b->h_hits, b->d_hits, (size_t)kept * sizeof(HitRec),
                          hipMemcpyDeviceToHost, b->stream));
    OG_TRY(hipStreamSynchronize(b->stream));
  }
  if (found > b->o.ring_cap) {
    put_err(err, errcap, "one step produced " + std::to_string(found) + " survivors and only " +
                             std::to_string(b->o.ring_cap) + " fit in the ring");
    return -2;
  }
  if (kept > cap) {
    put_err(err, errcap, "the output buffer holds " + std::to_string(cap) + " hits, " +
                             std::to_string(kept) + " were found");
    return -3;
  }
  for (uint32_t i = 0; i < kept; ++i) {
    const HitRec &h = b->h_hits[i];
    og_batch_hit &o = out[i];
    o.slot = h.slot;
    o.tag = h.slot < b->slots.size() ? b->slots[h.slot].tag : 0;
    o.offset = h.k * 8;
    for (int l = 0; l < 8; ++l)
      for (int k = 0; k < 4; ++k) o.packed[4 * l + k] = (uint8_t)(h.y[l] >> (8 * k));
    if (h.slot < b->slots.size()) b->slots[h.slot].hits++;
  }
  *n = kept;
  return 0;
}

int og_batch_slot_state(const og_batch *b, int slot, uint64_t *tag, uint64_t *examined,
                        uint64_t *hits, int *active) {
  if (!b || slot < 0 || (size_t)slot >= b->slots.size()) return -1;
  const HostSlot &s = b->slots[slot];
  if (tag) *tag = s.tag;
  if (examined) *examined = s.examined;
  if (hits) *hits = s.hits;
  if (active) *active = s.active ? 1 : 0;
  return 0;
}

uint64_t og_batch_per_step(const og_batch *b) { return b ? b->per_step : 0; }

double og_batch_last_ms(const og_batch *b) { return b ? b->last_ms : 0.0; }

void og_batch_destroy(og_batch *b) { free_batch(b); }

}  // extern "C"
