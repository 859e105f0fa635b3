// The serve loop: searches arrive over time and are served by one slot engine.
//
// Continuous batching: an arriving request becomes resident on the device at
// once (its chain state and filter tables are set up) when its footprint fits
// the device budget, else it waits in a bounded FIFO queue, and an arrival
// that finds the queue full is rejected (the back-pressure policy).  Every
// iteration runs one batch step, in which the slot engine's policy picks the
// resident requests that run (least served first by default), credits each
// survivor to its request, and retires a request the step its n-th hit
// arrives, which frees its memory for the queue.  A request records only the
// hits it asked for (og_batch_set_want), so easy filters cannot flood the ring.
// A request's chain state and filter tables stay on the device for its whole
// life; no request waits for another to finish.
//
// There is no serving layer in the reference; its nearest analogue is the
// master's lease queue (src/master/queue.rs:51-175).  Each request's search is
// exactly the single-search engine's: the same candidates, the same hits.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
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

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return -1.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

struct Live {
  size_t req;
  og_filterset *fs;
  size_t bytes;
};

}  // namespace

extern "C" {

int og_serve_run(int ordinal, const og_batch_opts *opts, const og_request *reqs, size_t n,
                 double duration_s, uint32_t max_queue, uint64_t device_budget,
                 og_request_result *results, og_serve_stats *stats, char *err, size_t errcap) {
  if (!reqs || !results || !stats) {
    put_err(err, errcap, "null argument");
    return -1;
  }
  for (size_t i = 1; i < n; ++i)
    if (reqs[i].arrival_s < reqs[i - 1].arrival_s) {
      put_err(err, errcap, "requests must be sorted by arrival time");
      return -1;
    }
  og_batch *b = nullptr;
  if (og_batch_create(ordinal, opts, &b, err, errcap) != 0) return -1;
  og_batch_opts o = opts ? *opts : og_batch_opts{};
  const uint32_t ring = o.ring_cap ? o.ring_cap : (1u << 16);
  const uint64_t budget = device_budget ? device_budget : (4ull << 30);
  std::vector<og_batch_hit> hits(ring);
  for (size_t i = 0; i < n; ++i) {
    results[i] = og_request_result{};
    results[i].admitted_s = results[i].first_hit_s = results[i].done_s = -1.0;
    results[i].status = OG_REQ_NOT_ARRIVED;
  }
  std::deque<size_t> queue;
  std::vector<Live> live;          // indexed by search id
  size_t resident = 0, max_resident = 0;
  uint64_t used = 0;
  size_t next = 0;
  uint64_t steps = 0, examined = 0;
  double busy_ms = 0.0;
  int rc = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&] {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  // admission: first come, first served, while the footprint fits the budget.
  // A request that could never fit (its footprint alone exceeds the budget),
  // whose filter does not parse, or that asks for no hit is rejected and
  // leaves the queue, so it never blocks the requests behind it.
  auto admit_pending = [&] {
    while (!queue.empty()) {
      const size_t r = queue.front();
      char e2[256];
      char text[sizeof reqs[r].filter + 1];
      const size_t len = strnlen(reqs[r].filter, sizeof reqs[r].filter);
      og_filterset *fs = nullptr;
      if (len < sizeof reqs[r].filter && reqs[r].want_hits != 0) {
        std::memcpy(text, reqs[r].filter, len + 1);
        fs = og_filters_parse(text, 0, e2, sizeof e2);
      }
      if (!fs) {
        results[r].status = OG_REQ_REJECTED;
        queue.pop_front();
        continue;
      }
      const size_t bytes = og_batch_footprint(b, fs);
      if (bytes > budget) {
        og_filters_free(fs);
        results[r].status = OG_REQ_REJECTED;
        queue.pop_front();
        continue;
      }
      if (used + bytes > budget) {
        og_filters_free(fs);
        break;  // wait for a resident request to free device memory
      }
      uint8_t start[128];
      og_key_start_point(reqs[r].seed, start);
      int id = -1;
      if (og_batch_admit(b, start, fs, (uint64_t)r, &id, e2, sizeof e2) != 0) {
        og_filters_free(fs);
        results[r].status = OG_REQ_REJECTED;
        queue.pop_front();
        continue;
      }
      queue.pop_front();
      // the request records only the hits it asked for
      (void)og_batch_set_want(b, id, reqs[r].want_hits);
      if ((size_t)id >= live.size()) live.resize(id + 1, Live{0, nullptr, 0});
      live[id] = Live{r, fs, bytes};
      used += bytes;
      resident++;
      max_resident = std::max(max_resident, resident);
      results[r].status = OG_REQ_RUNNING;
      results[r].admitted_s = now();
    }
  };
  for (;;) {
    const double t = now();
    if (t >= duration_s) break;
    while (next < n && reqs[next].arrival_s <= t) {
      // the arrival joins the FIFO and admission runs; only requests still
      // waiting after that count against the queue bound
      results[next].status = OG_REQ_QUEUED;
      queue.push_back(next);
      admit_pending();
      if (queue.size() > max_queue) {
        results[queue.back()].status = OG_REQ_REJECTED;
        queue.pop_back();
      }
      ++next;
    }
    admit_pending();
    if (resident == 0) {
      if (next >= n && queue.empty()) break;  // nothing resident, nothing left to arrive
      // nothing to run until the next arrival (admission has emptied the queue
      // or rejected its head, so the queue cannot wait on an empty device)
      const double wait =
          next < n ? std::min(reqs[next].arrival_s - now(), 0.001) : 0.001;
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    uint32_t nh = 0;
    if (og_batch_step(b, hits.data(), ring, &nh, err, errcap) != 0) {
      rc = -1;
      break;
    }
    steps++;
    busy_ms += og_batch_last_ms(b);
    const double t_step = now();
    examined = og_batch_examined(b);
    for (uint32_t i = 0; i < nh; ++i) {
      const og_batch_hit &h = hits[i];
      if (h.search >= live.size() || !live[h.search].fs) continue;
      og_request_result &res = results[live[h.search].req];
      if (res.hits == 0) res.first_hit_s = t_step;
      if (res.hits < 16) {
        res.found[res.hits].offset = h.offset;
        std::memcpy(res.found[res.hits].packed, h.packed, 32);
      }
      res.hits++;
    }
    // completion: retire a request the step its n-th hit arrives
    for (size_t id = 0; id < live.size(); ++id) {
      Live &l = live[id];
      if (!l.fs) continue;
      og_request_result &res = results[l.req];
      uint64_t ex = 0;
      og_batch_search_state(b, (int)id, nullptr, &ex, nullptr, nullptr);
      res.examined = ex;
      res.seen = og_batch_search_found(b, (int)id);
      if (res.hits >= reqs[l.req].want_hits) {
        res.done_s = t_step;
        res.status = OG_REQ_DONE;
        og_batch_retire(b, (int)id);
        og_filters_free(l.fs);
        used -= l.bytes;
        resident--;
        l = Live{0, nullptr, 0};
      }
    }
  }
  const double wall = now();
  for (Live &l : live)
    if (l.fs) og_filters_free(l.fs);
  og_batch_destroy(b);

  og_serve_stats st{};
  st.wall_s = wall;
  st.steps = steps;
  st.examined = examined;
  st.max_resident = (uint32_t)max_resident;
  st.keys_per_s = wall > 0 ? (double)examined / wall : 0.0;
  st.device_busy_s = busy_ms * 1e-3;
  std::vector<double> ttn, ttf;
  for (size_t i = 0; i < n; ++i) {
    og_request_result &r = results[i];
    switch (r.status) {
      case OG_REQ_DONE: st.completed++; break;
      case OG_REQ_RUNNING: st.running++; break;
      case OG_REQ_QUEUED: st.queued++; break;
      case OG_REQ_REJECTED: st.rejected++; break;
      default: st.not_arrived++; break;
    }
    if (r.status == OG_REQ_DONE) ttn.push_back(r.done_s - reqs[i].arrival_s);
    if (r.first_hit_s >= 0) ttf.push_back(r.first_hit_s - reqs[i].arrival_s);
  }
  st.p50_time_to_n = percentile(ttn, 0.50);
  st.p99_time_to_n = percentile(ttn, 0.99);
  st.p50_time_to_first = percentile(ttf, 0.50);
  st.p99_time_to_first = percentile(ttf, 0.99);
  *stats = st;
  return rc;
}

}  // extern "C"
