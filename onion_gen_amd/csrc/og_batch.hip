// The slot engine: many independent searches in one launch.
//
// A batch keeps any number of *resident* searches on the device - each with
// its own chain state (chains x 160 B in HBM), filter tables and next
// candidate index - and runs `slots` of them per step.  One step is a captured
// HIP graph: clear the survivor counters, run every chain of every scheduled
// search for `rounds` windows (two kernels, one per filter family, as parallel
// branches), copy the counters back.  Before each launch the
// host writes the per-slot descriptors (which search, its state, its filter,
// its k base), so scheduling never re-captures the graph.  With more resident
// searches than slots the device is time-sliced between them and no search
// waits for another to finish.  The default policy gives the slots to the
// searches that have examined the fewest candidates (least attained service):
// a new search runs at the next step, which keeps time-to-first-hit short when
// job sizes span many orders of magnitude; round-robin is the alternative.
// Searches of different filter kinds share the launch; each block reads its
// slot's descriptor once.
//
// This is the batching the reference does not have (its device path runs one
// search per engine, src/gpu/mod.rs:292-477); per search the arithmetic, the
// candidate numbering and the survivor contract are the single engine's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <ctime>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"
#include "og_devfilter.hpp"
#include "og_host.hpp"
#include "og_search.hpp"

using namespace og;

namespace {

// One unit slot of a step.  A search of width w takes w consecutive unit
// slots; unit i runs its chains [i*chains, (i+1)*chains).
struct SlotDesc {
  DevFilterArgs f;        // f.mode == 0: no search in this slot this step
  uint64_t k_base;        // first candidate of this step for the slot's search
  uint32_t *state;        // the search's chains: [4 coords][10 limbs][chains_total]
  const uint32_t *stride; // (y+x, y-x, 2dxy) of chains_total*(2H+1)*8G, Z = 1
  uint32_t budget;        // hit records the search may still write this step
  uint32_t chain_lo;      // the search's first chain in this unit
  uint32_t chains_total;  // the search's chains
  uint32_t lead;          // the search's first unit slot: its hit counter
};

struct BatchArgs {
  const uint32_t *table;
  uint32_t *scratch;    // [scratch_planes(half, true)][10][slots*chains]
  const SlotDesc *slots;
  uint32_t nslots, chains, half, rounds;  // chains: per unit slot
  uint32_t *counter;    // [0]: ring records; [1 + slot]: every hit of the search led there
  HitRec *hits;
  uint32_t cap;
  uint32_t *parked;     // one round per step: per column, the candidates left for
                        // the address kernel; null: drained in the search kernel
};

template <uint32_t MODE>
__device__ __forceinline__ void batch_body(const BatchArgs &a, const SlotDesc &d, uint32_t slot,
                                           uint32_t local, const ScanBank *banks,
                                           const uint16_t *dfa_lds) {
  const uint32_t chain = d.chain_lo + local;  // within the search
  Point10 P = ld_point(d.state, d.chains_total, chain);
  const cptr stride = (cptr)d.stride;
  Cached10 st;
  st.ypx = ld10_const(stride);
  st.ymx = ld10_const(stride + 10);
  st.t2d = ld10_const(stride + 20);
  HitSink sink{a.counter, a.hits, a.cap, slot, a.counter + 1 + d.lead, d.budget};
  search_chain<MODE>(P, st, (cptr)a.table, a.scratch, a.nslots * a.chains,
                     slot * a.chains + local, a.half, a.rounds, d.chains_total, chain, d.k_base,
                     d.f, banks, dfa_lds, a.parked, sink);
  st_point(d.state, d.chains_total, chain, P);
}

// One kernel per filter family, so that each gets its own register
// allocation: the key forms (prefix, bitmap) never carry the registers of the
// checksum path.  Both run over every slot; a block whose slot belongs to the
// other family (or is empty) returns at once.
template <bool kPatternFamily>
__global__ __launch_bounds__(256) OG_SEARCH_ATTR(kPatternFamily ? kModePattern : 0u) void
og_batch_kernel(BatchArgs a) {
  const uint32_t blocks_per_slot = a.chains / 256;
  const uint32_t slot = blockIdx.x / blocks_per_slot;
  const uint32_t local = (blockIdx.x % blocks_per_slot) * 256 + threadIdx.x;
  const SlotDesc d = a.slots[slot];  // block-uniform
  if (((d.f.mode & kModePattern) != 0) != kPatternFamily) return;
  constexpr uint32_t kAnyKey = kModePrefix | kModeBitmap;  // the key form read at run time
  if (kPatternFamily) {
    __shared__ ScanBank banks[kMaxBanks];
    __shared__ uint16_t dfa_lds[kLdsStates * 32];
    load_filter(banks, dfa_lds, d.f);
    batch_body<kModePattern | kAnyKey>(a, d, slot, local, banks, dfa_lds);
  } else if (d.f.mode) {
    batch_body<kAnyKey>(a, d, slot, local, nullptr, nullptr);
  }
}

// The deferred address stage of the pattern family's slots (a one-round
// step), at the occupancy the search loop's registers deny the search kernel.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void
og_batch_address_kernel(BatchArgs a) {
  const uint32_t blocks_per_slot = a.chains / 256;
  const uint32_t slot = blockIdx.x / blocks_per_slot;
  const uint32_t local = (blockIdx.x % blocks_per_slot) * 256 + threadIdx.x;
  const SlotDesc d = a.slots[slot];  // block-uniform
  if (!(d.f.mode & kModePattern)) return;
  __shared__ ScanBank banks[kMaxBanks];
  __shared__ uint16_t dfa_lds[2];  // the address stage walks its automata in global memory
  load_filter(banks, dfa_lds, d.f, false, false);
  HitSink sink{a.counter, a.hits, a.cap, slot, a.counter + 1 + d.lead, d.budget};
  const uint32_t g = slot * a.chains + local;
  address_stage(a.scratch, a.half, a.nslots * a.chains, g, a.parked[g], d.f, banks, sink);
}

// Clears the ring counter and the per-slot counts (a kernel rather than a
// memset node: the graph is replayed many times).
__global__ void og_batch_clear(uint32_t *counter, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counter[i] = 0;
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

struct Search {
  bool live = false;
  uint64_t tag = 0;
  uint64_t k_next = 0;    // next candidate index of this search
  uint64_t examined = 0;
  uint64_t hits = 0;      // hits recorded (written to the caller)
  uint64_t found = 0;     // hits found, recorded or not
  uint64_t want = ~0ull;  // record at most this many hits in all
  uint64_t seq = 0;       // admission order
  uint32_t width = 0;     // unit slots it takes; chains = width * unit chains
  uint8_t start[128];     // its s0*B, for laying the chains out again
  uint32_t *d_state = nullptr;
  DevFilterHost filt;
  size_t bytes = 0;       // device bytes held (state + tables)
};

}  // namespace

struct og_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;         // the pattern family's branch of the step
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  og_batch_opts o{};
  uint64_t per_step = 0;  // candidates one step examines per scheduled search
  uint32_t *d_table = nullptr, *d_scratch = nullptr, *d_parked = nullptr;
  SlotDesc *d_slots = nullptr, *h_slots = nullptr;
  uint32_t *d_counter = nullptr;
  HitRec *d_hits = nullptr;
  uint32_t *h_counter = nullptr;
  HitRec *h_hits = nullptr;
  std::vector<Search> searches;       // indexed by search id
  std::vector<uint32_t> free_ids;
  std::deque<uint32_t> runnable;      // resident searches, in round-robin order
  std::vector<uint32_t> pick;         // this step's choice
  std::vector<uint32_t> lead;         // and the first unit slot of each
  uint64_t admitted = 0;              // admission counter
  std::vector<int64_t> assigned;      // search id per slot in the last step, -1 empty
  std::vector<uint32_t *> d_stride;   // per width: the stride point of its chains
  uint64_t examined = 0;              // candidates examined by every step
  BatchArgs args;
  HostCurve curve;
  float last_ms = 0;
  uint64_t steps = 0;
  size_t resident_bytes = 0;
};

namespace {

void drop_search(Search &s) {
  (void)hipFree(s.d_state);
  s.d_state = nullptr;
  s.filt.free_device();
  s.live = false;
}

void free_batch(og_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (Search &s : b->searches)
    if (s.live) drop_search(s);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (uint32_t *p : b->d_stride) (void)hipFree(p);
  (void)hipFree(b->d_table);
  (void)hipFree(b->d_scratch);
  (void)hipFree(b->d_parked);
  (void)hipFree(b->d_slots);
  (void)hipFree(b->d_counter);
  (void)hipFree(b->d_hits);
  (void)hipHostFree(b->h_slots);
  (void)hipHostFree(b->h_counter);
  (void)hipHostFree(b->h_hits);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// The stride point of a search of `width` unit slots: every chain advances by
// chains_total * (2H+1) * 8G per window round.  Built on first use.
hipError_t stride_for(og_batch *b, uint32_t width, const uint32_t **out) {
  if (!b->d_stride[width]) {
    const uint64_t S = 2ull * b->o.half + 1;
    const Cached1 c = b->curve.multiple_of_eight_cached1((uint64_t)width * b->o.chains * S);
    uint32_t words[30];
    put_fe10(words, c.ypx);
    put_fe10(words + 10, c.ymx);
    put_fe10(words + 20, c.t2d);
    hipError_t e = hipMalloc(&b->d_stride[width], sizeof words);
    if (e == hipSuccess)
      e = hipMemcpy(b->d_stride[width], words, sizeof words, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(b->d_stride[width]);
      b->d_stride[width] = nullptr;
      return e;
    }
  }
  *out = b->d_stride[width];
  return hipSuccess;
}

// Lays a search's chains out again at `width` unit slots, from its next
// candidate: chain c's first window starts at k_next + c(2H+1).  Its examined
// candidates stay the prefix [0, k_next), so nothing is skipped or repeated.
hipError_t layout(og_batch *b, Search &s, uint32_t width) {
  const uint32_t chains = width * b->o.chains;
  const uint32_t *stride = nullptr;
  hipError_t e = stride_for(b, width, &stride);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(b->stream);  // no step still reads the old state
  if (e != hipSuccess) return e;
  uint32_t *state = nullptr;
  e = hipMalloc(&state, (size_t)chains * 160);
  if (e != hipSuccess) return e;
  SetupArgs sa = b->curve.setup_args(s.start, state, chains, 0, chains, b->o.half, s.k_next);
  hipLaunchKernelGGL(og_batch_setup, dim3(chains / 256), dim3(256), 0, b->stream, sa);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    (void)hipFree(state);
    return e;
  }
  (void)hipFree(s.d_state);
  s.d_state = state;
  b->resident_bytes -= s.bytes;
  s.bytes = (size_t)chains * 160 + s.filt.device_bytes();
  b->resident_bytes += s.bytes;
  s.width = width;
  return hipSuccess;
}

// The width every resident search should have: the unit slots shared out.
uint32_t target_width(const og_batch *b, size_t live) {
  if (live == 0) return 1;
  return (uint32_t)std::max<size_t>(1, std::min<size_t>(b->o.slots, b->o.slots / live));
}

// The cache key of the slot engine's unit shape on this device:
// device/<name>/hip-batch/0/<version> (src/tuning.rs:65-74's key shape).
std::string batch_tuning_key(int ordinal) {
  char name[256] = {0};
  uint64_t mem = 0;
  int cus = 0;
  og_hip_device_info(ordinal, name, sizeof name, &mem, &cus);
  std::string key(og_tuning_device_key(name, "hip-batch", 0, nullptr, 0), '\0');
  og_tuning_device_key(name, "hip-batch", 0, &key[0], key.size() + 1);
  return key;
}

// Device bytes of a slot engine of this shape with every unit slot taken:
// the scratch planes, the offset table, the ring, and the chain state.
uint64_t batch_bytes(uint32_t slots, uint32_t chains, uint32_t half, uint32_t ring) {
  const uint64_t T = (uint64_t)slots * chains;
  return scratch_planes(half, true) * 40 * T + (uint64_t)half * kTableWords * 4 +
         (uint64_t)ring * sizeof(HitRec) + 160 * T;
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
  if (!o.slots && !o.chains && !o.half) {
    // the unit shape the tuning walk recorded for this device, if any
    og_tuning_shape t{};
    if (og_tuning_get(batch_tuning_key(ordinal).c_str(), &t) == 1 && t.slots && t.second &&
        t.first) {
      o.slots = t.slots;
      o.chains = t.second;
      o.half = t.first;
    }
  }
  if (!o.slots) o.slots = 64;
  if (!o.chains) o.chains = 16384;
  if (!o.half) o.half = 256;
  if (!o.rounds) o.rounds = 1;
  if (!o.ring_cap) o.ring_cap = 1u << 16;
  if (o.chains % 256 || (uint64_t)o.slots * o.chains > (1u << 22) || o.half > 4096 ||
      o.slots > 4096) {
    put_err(err, errcap,
            "chains must be a multiple of 256, slots*chains at most 2^22, half at most 4096");
    return -1;
  }
  if (o.policy > OG_POLICY_ROUND_ROBIN) {
    put_err(err, errcap, "unknown scheduling policy " + std::to_string(o.policy));
    return -1;
  }
  std::unique_ptr<og_batch, void (*)(og_batch *)> b(new og_batch(), free_batch);
  b->device = ordinal;
  b->o = o;
  b->assigned.assign(o.slots, -1);
  const uint64_t S = 2ull * o.half + 1;
  b->per_step = (uint64_t)o.chains * S * o.rounds;
  const size_t T = (size_t)o.slots * o.chains;
  OG_TRY(hipSetDevice(ordinal));
  OG_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  OG_TRY(hipEventCreate(&b->ev0));
  OG_TRY(hipEventCreate(&b->ev1));
  OG_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  OG_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  OG_TRY(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
  const std::vector<uint32_t> table = b->curve.offset_table10(o.half);
  OG_TRY(hipMalloc(&b->d_table, std::max<size_t>(table.size(), kTableWords) * 4));
  OG_TRY(hipMalloc(&b->d_scratch, scratch_planes(o.half, true) * 10 * T * 4));
  OG_TRY(hipMalloc(&b->d_slots, o.slots * sizeof(SlotDesc)));
  OG_TRY(hipHostMalloc(&b->h_slots, o.slots * sizeof(SlotDesc), hipHostMallocDefault));
  OG_TRY(hipMalloc(&b->d_counter, (1 + (size_t)o.slots) * 4));
  OG_TRY(hipMalloc(&b->d_hits, (size_t)o.ring_cap * sizeof(HitRec)));
  OG_TRY(hipHostMalloc(&b->h_counter, (1 + (size_t)o.slots) * 4, hipHostMallocDefault));
  OG_TRY(hipHostMalloc(&b->h_hits, (size_t)o.ring_cap * sizeof(HitRec), hipHostMallocDefault));
  if (!table.empty())
    OG_TRY(hipMemcpyAsync(b->d_table, table.data(), table.size() * 4, hipMemcpyHostToDevice,
                          b->stream));
  std::memset(b->h_slots, 0, o.slots * sizeof(SlotDesc));
  OG_TRY(hipMemcpyAsync(b->d_slots, b->h_slots, o.slots * sizeof(SlotDesc),
                        hipMemcpyHostToDevice, b->stream));
  OG_TRY(hipStreamSynchronize(b->stream));

  BatchArgs &a = b->args;
  std::memset(&a, 0, sizeof a);
  a.table = b->d_table;
  a.scratch = b->d_scratch;
  a.slots = b->d_slots;
  a.nslots = o.slots;
  a.chains = o.chains;
  a.half = o.half;
  a.rounds = o.rounds;
  b->d_stride.assign(o.slots + 1, nullptr);
  a.counter = b->d_counter;
  a.hits = b->d_hits;
  a.cap = o.ring_cap;
  if (o.rounds == 1) {
    // one round per step: the pattern family's address stage gets its own kernel
    OG_TRY(hipMalloc(&b->d_parked, (size_t)T * 4));
    a.parked = b->d_parked;
  }

  // Capture one step: clear, then the two family kernels as parallel branches
  // (fork/join through events; the pattern branch followed by its address
  // kernel), then the counters back to the host.
  OG_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(og_batch_clear, dim3(1), dim3(256), 0, b->stream, b->d_counter,
                     1u + o.slots);
  (void)hipEventRecord(b->fork, b->stream);
  (void)hipStreamWaitEvent(b->side, b->fork, 0);
  hipLaunchKernelGGL(og_batch_kernel<false>, dim3((unsigned)(T / 256)), dim3(256), 0, b->stream,
                     a);
  hipLaunchKernelGGL(og_batch_kernel<true>, dim3((unsigned)(T / 256)), dim3(256), 0, b->side, a);
  if (a.parked)
    hipLaunchKernelGGL(og_batch_address_kernel, dim3((unsigned)(T / 256)), dim3(256), 0, b->side,
                       a);
  (void)hipEventRecord(b->join, b->side);
  (void)hipStreamWaitEvent(b->stream, b->join, 0);
  (void)hipMemcpyAsync(b->h_counter, b->d_counter, (1 + (size_t)o.slots) * 4,
                       hipMemcpyDeviceToHost, b->stream);
  OG_TRY(hipStreamEndCapture(b->stream, &b->graph));
  OG_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  *out = b.release();
  return 0;
}

int og_batch_tune(int ordinal, const og_filterset *const *filters, size_t nfilters,
                  uint64_t max_bytes, double max_step_ms, og_batch_opts *best, double *best_rate,
                  char *err, size_t errcap) {
  // a step bounds how soon a new search runs and how soon a found hit is
  // reported, so only shapes whose steps stay within the bound are eligible
  const double step_cap = max_step_ms > 0 ? max_step_ms : 25.0;
  if (!filters || nfilters == 0) {
    put_err(err, errcap, "the walk needs at least one filter set");
    return -1;
  }
  for (size_t i = 0; i < nfilters; ++i)
    if (!filterset_of(filters[i])) {
      put_err(err, errcap, "null filter set");
      return -1;
    }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || ordinal < 0 || ordinal >= count) {
    put_err(err, errcap, "no such HIP device");
    return -1;
  }
  uint64_t budget = max_bytes;
  if (!budget) {
    size_t free_b = 0, total_b = 0;
    OG_TRY(hipSetDevice(ordinal));
    OG_TRY(hipMemGetInfo(&free_b, &total_b));
    budget = free_b / 2;
  }
  // unit slots x chains per slot x pairs per window, every search on one slot
  std::vector<og_batch_opts> walk;
  for (uint32_t slots : {32u, 64u, 128u})
    for (uint32_t chains : {4096u, 8192u, 16384u, 32768u})
      for (uint32_t half : {256u, 512u}) {
        if ((uint64_t)slots * chains > (1u << 22)) continue;
        og_batch_opts o{};
        o.slots = slots;
        o.chains = chains;
        o.half = half;
        o.rounds = 1;
        o.ring_cap = 1u << 16;
        o.policy = OG_POLICY_LEAST_SERVED;
        walk.push_back(o);
      }
  std::vector<double> rates(walk.size(), 0.0);
  std::vector<uint64_t> sizes(walk.size(), 0);
  std::string last = "no shape was tried";
  for (size_t w = 0; w < walk.size(); ++w) {
    const og_batch_opts &o = walk[w];
    sizes[w] = batch_bytes(o.slots, o.chains, o.half, o.ring_cap);
    if (sizes[w] > budget) {
      last = "every shape exceeds the memory budget of " + std::to_string(budget) + " bytes";
      continue;
    }
    og_batch *b = nullptr;
    char e2[512] = {0};
    if (og_batch_create(ordinal, &o, &b, e2, sizeof e2) != 0) {
      last = e2;
      continue;
    }
    // as many searches as slots, the given filter sets in turn, each on its
    // own block's chain; each records one hit, so no set floods the ring
    bool ok = true;
    for (uint32_t i = 0; i < o.slots && ok; ++i) {
      uint8_t root[32] = {0}, seed[32], start[128];
      std::memcpy(root, "slot engine tuning walk", 23);
      og_run_block_seed(root, i, seed);
      og_key_start_point(seed, start);
      int id = -1;
      ok = og_batch_admit(b, start, filters[i % nfilters], i, &id, e2, sizeof e2) == 0 &&
           og_batch_set_want(b, id, 1) == 0;
    }
    std::vector<og_batch_hit> hits(o.ring_cap);
    uint32_t n = 0;
    double rate = 0;
    // one step warms up; then steps for at least 0.25 s (and three)
    if (ok && og_batch_step(b, hits.data(), o.ring_cap, &n, e2, sizeof e2) == 0) {
      const uint64_t ex0 = og_batch_examined(b);
      const auto t0 = std::chrono::steady_clock::now();
      double sec = 0;
      int steps = 0;
      while (ok && (steps < 3 || sec < 0.25)) {
        ok = og_batch_step(b, hits.data(), o.ring_cap, &n, e2, sizeof e2) == 0;
        ++steps;
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
      if (ok && sec > 0) rate = (double)(og_batch_examined(b) - ex0) / sec;
      if (ok && steps > 0 && sec * 1e3 / steps > step_cap) {
        rate = 0;
        last = "a step takes " + std::to_string(sec * 1e3 / steps) + " ms, over the bound of " +
               std::to_string(step_cap) + " ms";
      }
    }
    if (rate == 0 && e2[0]) last = e2;
    og_batch_destroy(b);
    rates[w] = rate;
  }
  const size_t pick = og_tuning_choose(rates.data(), sizes.data(), walk.size(), budget, 0.03);
  if (pick >= walk.size()) {
    put_err(err, errcap, "no unit shape this device would take: " + last);
    return -1;
  }
  og_tuning_shape t{};
  t.first = walk[pick].half;
  t.second = walk[pick].chains;
  t.slots = walk[pick].slots;
  t.rounds = 1;
  t.rate = rates[pick];
  t.at = (int64_t)std::time(nullptr);
  t.radix = 10;
  og_tuning_put(batch_tuning_key(ordinal).c_str(), &t);
  if (best) *best = walk[pick];
  if (best_rate) *best_rate = rates[pick];
  return 0;
}

size_t og_batch_footprint(const og_batch *b, const og_filterset *filters) {
  const FilterSet *fs = filterset_of(filters);
  if (!fs) return 0;
  DevFilterHost h;
  if (!h.compile(*fs).empty()) return 0;
  const size_t chains = b ? b->o.chains : 2048;
  return h.device_bytes() + chains * 160;
}

int og_batch_admit(og_batch *b, const uint8_t start_xyzt[128], const og_filterset *filters,
                   uint64_t tag, int *id_out, char *err, size_t errcap) {
  if (!b || !start_xyzt || !filters || !id_out) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  uint32_t id;
  if (!b->free_ids.empty()) {
    id = b->free_ids.back();
    b->free_ids.pop_back();
  } else {
    id = (uint32_t)b->searches.size();
    b->searches.emplace_back();
  }
  Search &s = b->searches[id];
  s = Search();
  std::string msg = s.filt.compile(*filterset_of(filters));
  if (!msg.empty()) {
    b->free_ids.push_back(id);
    put_err(err, errcap, msg);
    return -1;
  }
  OG_TRY(hipSetDevice(b->device));
  std::memcpy(s.start, start_xyzt, 128);
  hipError_t e = s.filt.to_device(b->stream);
  if (e == hipSuccess) e = layout(b, s, target_width(b, b->runnable.size() + 1));
  if (e != hipSuccess) {
    b->resident_bytes -= s.bytes;
    s.bytes = 0;
    drop_search(s);
    b->free_ids.push_back(id);
    put_err(err, errcap, std::string("admission: ") + hipGetErrorString(e));
    return -1;
  }
  s.live = true;
  s.tag = tag;
  s.seq = b->admitted++;
  b->runnable.push_back(id);
  *id_out = (int)id;
  return 0;
}

int og_batch_retire(og_batch *b, int id) {
  if (!b || id < 0 || (size_t)id >= b->searches.size()) return -1;
  Search &s = b->searches[id];
  if (!s.live) return 0;
  (void)hipSetDevice(b->device);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
  b->resident_bytes -= s.bytes;
  drop_search(s);
  b->runnable.erase(std::remove(b->runnable.begin(), b->runnable.end(), (uint32_t)id),
                    b->runnable.end());
  for (int64_t &a : b->assigned)
    if (a == id) a = -1;
  b->free_ids.push_back((uint32_t)id);
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
  // Widths follow the number of resident searches: with fewer searches than
  // unit slots each is laid out wider (more chains from its next candidate),
  // so a lone search still fills the device; under contention wide searches
  // narrow again.  A search widens only when that gains a quarter or more.
  const uint32_t wt = target_width(b, b->runnable.size());
  uint64_t demand = 0;
  for (uint32_t id : b->runnable) demand += b->searches[id].width;
  for (uint32_t id : b->runnable) {
    Search &sr = b->searches[id];
    const bool widen = sr.width < wt && 4 * sr.width <= 3 * wt;
    const bool narrow = sr.width > wt && demand > b->o.slots;
    if (!widen && !narrow) continue;
    demand -= sr.width;
    OG_TRY(layout(b, sr, wt));
    demand += sr.width;
  }
  // The policy orders the resident searches; they take unit slots in that
  // order while their widths fit.
  std::vector<uint32_t> &pick = b->pick;
  pick.clear();
  if (b->o.policy == OG_POLICY_ROUND_ROBIN) {
    uint32_t used = 0;
    const size_t n_live = b->runnable.size();
    for (size_t i = 0; i < n_live && used < b->o.slots; ++i) {
      const uint32_t id = b->runnable.front();
      if (used + b->searches[id].width > b->o.slots) break;
      used += b->searches[id].width;
      pick.push_back(id);
      b->runnable.pop_front();
      b->runnable.push_back(id);
    }
  } else {
    std::vector<uint32_t> order(b->runnable.begin(), b->runnable.end());
    auto served_less = [b](uint32_t x, uint32_t y) {
      const Search &p = b->searches[x], &q = b->searches[y];
      return p.examined != q.examined ? p.examined < q.examined : p.seq < q.seq;
    };
    const size_t head = std::min<size_t>(order.size(), b->o.slots);
    std::partial_sort(order.begin(), order.begin() + head, order.end(), served_less);
    uint32_t used = 0;
    for (size_t i = 0; i < head; ++i) {
      const uint32_t w = b->searches[order[i]].width;
      if (used + w > b->o.slots) continue;
      used += w;
      pick.push_back(order[i]);
    }
  }
  std::memset(b->h_slots, 0, b->o.slots * sizeof(SlotDesc));
  for (size_t s = 0; s < b->o.slots; ++s) b->assigned[s] = -1;
  b->lead.clear();
  uint32_t unit = 0;
  for (uint32_t id : pick) {
    Search &sr = b->searches[id];
    const uint32_t *stride = nullptr;
    OG_TRY(stride_for(b, sr.width, &stride));
    const uint64_t left = sr.want > sr.hits ? sr.want - sr.hits : 0;
    b->lead.push_back(unit);
    for (uint32_t i = 0; i < sr.width; ++i, ++unit) {
      SlotDesc &d = b->h_slots[unit];
      b->assigned[unit] = id;
      d.f = sr.filt.args();
      d.f.hash = nullptr;  // the slot kernels keep the automaton path
      d.f.hash_key = d.f.hash_addr = 0;
      d.k_base = sr.k_next;
      d.state = sr.d_state;
      d.stride = stride;
      d.budget = (uint32_t)std::min<uint64_t>(left, 0xffffffffu);
      d.chain_lo = i * b->o.chains;
      d.chains_total = sr.width * b->o.chains;
      d.lead = b->lead.back();
    }
  }
  OG_TRY(hipMemcpyAsync(b->d_slots, b->h_slots, b->o.slots * sizeof(SlotDesc),
                        hipMemcpyHostToDevice, b->stream));
  OG_TRY(hipEventRecord(b->ev0, b->stream));
  OG_TRY(hipGraphLaunch(b->exec, b->stream));
  OG_TRY(hipEventRecord(b->ev1, b->stream));
  OG_TRY(hipStreamSynchronize(b->stream));
  OG_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  b->steps++;
  for (size_t i = 0; i < pick.size(); ++i) {
    Search &sr = b->searches[pick[i]];
    const uint64_t ran = (uint64_t)sr.width * b->per_step;
    sr.k_next += ran;
    sr.examined += ran;
    b->examined += ran;
    sr.found += b->h_counter[1 + b->lead[i]];
  }
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
  uint32_t m = 0;
  for (uint32_t i = 0; i < kept; ++i) {
    const HitRec &h = b->h_hits[i];
    if (h.slot >= b->o.slots || b->assigned[h.slot] < 0) continue;
    Search &sr = b->searches[b->assigned[h.slot]];
    og_batch_hit &o = out[m++];
    o.search = (uint32_t)b->assigned[h.slot];
    o.tag = sr.tag;
    o.offset = h.k * 8;
    for (int l = 0; l < 8; ++l)
      for (int k = 0; k < 4; ++k) o.packed[4 * l + k] = (uint8_t)(h.y[l] >> (8 * k));
    sr.hits++;
  }
  *n = m;
  return 0;
}

int og_batch_search_state(const og_batch *b, int id, uint64_t *tag, uint64_t *examined,
                          uint64_t *hits, int *live) {
  if (!b || id < 0 || (size_t)id >= b->searches.size()) return -1;
  const Search &s = b->searches[id];
  if (tag) *tag = s.tag;
  if (examined) *examined = s.examined;
  if (hits) *hits = s.hits;
  if (live) *live = s.live ? 1 : 0;
  return 0;
}

int og_batch_set_want(og_batch *b, int id, uint64_t want) {
  if (!b || id < 0 || (size_t)id >= b->searches.size() || !b->searches[id].live) return -1;
  b->searches[id].want = want;
  return 0;
}

uint64_t og_batch_search_found(const og_batch *b, int id) {
  if (!b || id < 0 || (size_t)id >= b->searches.size()) return 0;
  return b->searches[id].found;
}

uint32_t og_batch_search_width(const og_batch *b, int id) {
  if (!b || id < 0 || (size_t)id >= b->searches.size() || !b->searches[id].live) return 0;
  return b->searches[id].width;
}

uint64_t og_batch_examined(const og_batch *b) { return b ? b->examined : 0; }

uint64_t og_batch_per_step(const og_batch *b) { return b ? b->per_step : 0; }

uint64_t og_batch_resident_bytes(const og_batch *b) { return b ? b->resident_bytes : 0; }

double og_batch_last_ms(const og_batch *b) { return b ? b->last_ms : 0.0; }

void og_batch_destroy(og_batch *b) { free_batch(b); }

}  // extern "C"
