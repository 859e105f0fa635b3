// The lease adapter: a proto::Lease (src/proto.rs:47-57) searched by one
// device engine, finds reported as proto::Found (src/proto.rs:78-84).
//
// Block b of a lease is the chain of block_seed(root, b) (src/run.rs:211-220),
// the lease's own numbering kept (src/worker/run.rs:183-189), so the master
// derives every find's key from (block, offset) exactly as it would for a CPU
// worker.  A block is BATCHES_PER_BLOCK launches (src/run.rs:29) unless the
// caller says otherwise.  Every survivor goes through the batched hit sink
// first: a find is reported only when the key re-derived from its secret
// matches what the device saw (src/run.rs:1089-1104), and the filter set
// confirms it when one is given.
#include <cstring>
#include <string>
#include <vector>

#include "../../include/onion_gen_hip.h"

namespace {

void put_err(char *err, size_t cap, const std::string &msg) {
  if (err && cap) {
    size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

}  // namespace

extern "C" int og_hip_run_lease(og_engine *e, const og_lease *lease, const og_filterset *filters,
                                uint32_t launches, og_lease_find *out, size_t cap, size_t *n,
                                uint64_t *candidates, uint64_t *blocks_done, char *err,
                                size_t errcap) {
  if (!e || !lease || !n || (cap && !out)) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  *n = 0;
  if (candidates) *candidates = 0;
  if (blocks_done) *blocks_done = 0;
  const uint32_t per_block = launches ? launches : 64;
  const uint32_t room = og_hip_engine_room(e);
  std::vector<og_survivor> surv(room);
  std::vector<og_survivor> block_hits;
  std::vector<og_key_record> recs;
  for (uint64_t b = lease->first_block; b < lease->first_block + lease->blocks; ++b) {
    uint8_t seed[32], start[128];
    og_run_block_seed(lease->seed, b, seed);
    og_key_start_point(seed, start);
    if (og_hip_engine_restart(e, start, err, errcap) != 0) return -1;
    block_hits.clear();
    for (uint32_t l = 0; l < per_block; ++l) {
      uint32_t got = 0;
      if (og_hip_engine_advance(e, surv.data(), room, &got, err, errcap) != 0) return -1;
      block_hits.insert(block_hits.end(), surv.begin(), surv.begin() + got);
      if (candidates) *candidates += og_hip_engine_batch(e);
    }
    recs.resize(block_hits.size());
    if (og_hip_finish_hits(og_hip_engine_device(e), seed, block_hits.data(), block_hits.size(),
                           recs.data(), err, errcap) != 0)
      return -1;
    for (const og_key_record &r : recs) {
      if (r.status != OG_HIT_OK) continue;
      // a bitmap-only engine's survivors are "maybe": the set decides
      // (src/run.rs:1082 re-checks with FilterSet::match_key)
      if (filters) {
        uint8_t addr[35];
        og_key_address_bytes(r.public_key, addr);
        if (og_filters_match_address(filters, addr) < 0) continue;
      }
      if (*n >= cap) {
        put_err(err, errcap, "more finds than the output holds (" + std::to_string(cap) + ")");
        return -2;
      }
      out[(*n)++] = og_lease_find{lease->lease, b, r.offset};
    }
    if (blocks_done) ++*blocks_done;
  }
  return 0;
}
