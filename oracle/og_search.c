/*
 * CPU oracle, search driver.  TEST INFRASTRUCTURE ONLY (see og_oracle.h).
 *
 * The hit definition follows src/run.rs:453-598 (scan_batch) and
 * src/run.rs:740-776 (record_hit): a candidate k is a hit when its address
 * (src/key.rs:116) matches some filter of the set.  Candidate k is the point
 * (s0 + 8k)B, k = 0 being the start point itself (src/batch.rs:74-77).
 */
#include "og_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int ofilter_all_fast(const ofilterset *s);

static void advance_to(opoint *acc, uint64_t k0) {
  /* acc += k0 * 8G by double-and-add on the 64-bit multiple */
  if (!k0) return;
  ocached step;
  ocurve_eight_cached(&step);
  opoint add;
  uint8_t sc[32] = {0};
  uint64_t d = k0;
  /* k0*8 may exceed 64 bits; use a 128-bit scalar */
  unsigned __int128 s = (unsigned __int128)d * 8;
  for (int i = 0; i < 16; i++) sc[i] = (uint8_t)(s >> (8 * i));
  ocurve_scalar_base_mult(&add, sc);
  ocached c;
  ocurve_to_cached(&c, &add);
  ocurve_add(acc, acc, &c);
  (void)step;
}

size_t osearch(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
               const ofilterset *s, ohit *hits, size_t cap) {
  opoint acc;
  opoint_from_bytes(&acc, start_xyzt);
  advance_to(&acc, k0);
  obatch *e = obatch_new(width);
  int fast = ofilter_all_fast(s);
  size_t found = 0;
  for (uint64_t done = 0; done < n; done += width) {
    obatch_run(e, &acc);
    uint64_t take = n - done < width ? n - done : width;
    for (uint64_t i = 0; i < take; i++) {
      int which;
      uint8_t pk[32], ab[35];
      if (fast) {
        /* literal prefixes read only key symbols below 49: the sign is not needed */
        which = ofilter_match_key(s, obatch_packed(e, i));
        if (which < 0) continue;
        obatch_packed_with_sign(e, i, pk);
        okey_address_bytes(pk, ab);
      } else {
        obatch_packed_with_sign(e, i, pk);
        okey_address_bytes(pk, ab);
        which = ofilter_match_address(s, ab);
        if (which < 0) continue;
      }
      if (found < cap) {
        ohit *h = &hits[found];
        h->k = k0 + done + i;
        memcpy(h->pk, pk, 32);
        obase32_encode(ab, 35, h->address);
        h->filter = which;
      }
      found++;
    }
  }
  obatch_free(e);
  return found;
}

typedef struct {
  const uint8_t *start;
  uint64_t k0, n;
  size_t width;
  const ofilterset *s;
  ohit *hits;
  size_t cap, found;
} search_part;

static void *search_part_run(void *p) {
  search_part *a = p;
  a->found = osearch(a->start, a->k0, a->n, a->width, a->s, a->hits, a->cap);
  return NULL;
}

/* osearch over `threads` contiguous parts of [k0, k0+n).  A candidate's key
 * does not depend on where its batch starts (every canonical y is exact), so
 * the parts' hits, concatenated in order, are osearch's over the whole range. */
size_t osearch_mt(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
                  const ofilterset *s, int threads, ohit *hits, size_t cap) {
  if (threads < 1) threads = 1;
  search_part *parts = calloc((size_t)threads, sizeof *parts);
  pthread_t *th = calloc((size_t)threads, sizeof *th);
  const uint64_t chunk = (n / (uint64_t)threads + width - 1) / width * width;
  for (int t = 0; t < threads; t++) {
    const uint64_t lo = chunk * (uint64_t)t < n ? chunk * (uint64_t)t : n;
    const uint64_t hi = lo + chunk < n ? lo + chunk : n;
    parts[t].start = start_xyzt;
    parts[t].k0 = k0 + lo;
    parts[t].n = hi - lo;
    parts[t].width = width;
    parts[t].s = s;
    parts[t].cap = cap;
    parts[t].hits = malloc((cap ? cap : 1) * sizeof(ohit));
    pthread_create(&th[t], NULL, search_part_run, &parts[t]);
  }
  size_t found = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    const size_t got = parts[t].found < cap ? parts[t].found : cap;
    for (size_t i = 0; i < got && found + i < cap; i++) hits[found + i] = parts[t].hits[i];
    found += parts[t].found;
    free(parts[t].hits);
  }
  free(parts);
  free(th);
  return found;
}

void odigest(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
             uint64_t out[2]) {
  opoint acc;
  opoint_from_bytes(&acc, start_xyzt);
  advance_to(&acc, k0);
  obatch *e = obatch_new(width);
  uint64_t x = 0, sum = 0;
  for (uint64_t done = 0; done < n; done += width) {
    obatch_run(e, &acc);
    uint64_t take = n - done < width ? n - done : width;
    for (uint64_t i = 0; i < take; i++) {
      const uint8_t *p = obatch_packed(e, i);
      uint64_t w = 0;
      for (int b = 7; b >= 0; b--) w = (w << 8) | p[b];
      x ^= w;
      sum += w;
    }
  }
  obatch_free(e);
  out[0] = x;
  out[1] = sum;
}

typedef struct {
  uint8_t start[128];
  size_t width;
  double seconds;
  const ofilterset *s;
  uint64_t examined;
  uint64_t hits;
} bench_arg;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void *bench_thread(void *p) {
  bench_arg *a = p;
  opoint acc;
  opoint_from_bytes(&acc, a->start);
  obatch *e = obatch_new(a->width);
  double t0 = now_s();
  while (now_s() - t0 < a->seconds) {
    obatch_run(e, &acc);
    for (size_t i = 0; i < a->width; i++)
      if (ofilter_match_key(a->s, obatch_packed(e, i)) >= 0) a->hits++;
    a->examined += a->width;
  }
  obatch_free(e);
  return NULL;
}

double obench_batch(const uint8_t seed[32], size_t width, int threads, double seconds,
                    const ofilterset *s, uint64_t *examined) {
  pthread_t *th = calloc((size_t)threads, sizeof *th);
  bench_arg *args = calloc((size_t)threads, sizeof *args);
  double t0 = now_s();
  for (int t = 0; t < threads; t++) {
    /* one independent chain per thread, like the reference's one block per worker */
    uint8_t bs[32];
    orun_block_seed(seed, (uint64_t)t, bs);
    okey_start_point(bs, args[t].start);
    args[t].width = width;
    args[t].seconds = seconds;
    args[t].s = s;
    pthread_create(&th[t], NULL, bench_thread, &args[t]);
  }
  uint64_t total = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    total += args[t].examined;
  }
  double dt = now_s() - t0;
  free(th);
  free(args);
  if (examined) *examined = total;
  return dt > 0 ? total / dt : 0;
}

typedef struct {
  uint8_t start[128];
  size_t batch;
  double seconds;
  const ofilterset *s;
  uint64_t examined;
  uint64_t hits;
} pairs_arg;

static void *pairs_thread(void *p) {
  pairs_arg *a = p;
  opoint acc;
  opoint_from_bytes(&acc, a->start);
  opairs *e = opairs_new(a->batch);
  opairs_begin(e, &acc);
  const size_t n = opairs_batch_size(e);
  double t0 = now_s();
  while (now_s() - t0 < a->seconds) {
    const size_t k = opairs_run(e, &acc, a->s);
    for (size_t i = 0; i < k; i++) {
      uint32_t idx;
      uint8_t packed[32];
      opairs_survivor(e, i, &idx, packed);
      if (ofilter_match_key(a->s, packed) >= 0) a->hits++;
    }
    a->examined += n;
  }
  opairs_free(e);
  return NULL;
}

double obench_pairs(const uint8_t seed[32], size_t batch, int threads, double seconds,
                    const ofilterset *s, uint64_t *examined) {
  pthread_t *th = calloc((size_t)threads, sizeof *th);
  pairs_arg *args = calloc((size_t)threads, sizeof *args);
  double t0 = now_s();
  for (int t = 0; t < threads; t++) {
    uint8_t bs[32];
    orun_block_seed(seed, (uint64_t)t, bs);
    okey_start_point(bs, args[t].start);
    args[t].batch = batch;
    args[t].seconds = seconds;
    args[t].s = s;
    pthread_create(&th[t], NULL, pairs_thread, &args[t]);
  }
  uint64_t total = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    total += args[t].examined;
  }
  double dt = now_s() - t0;
  free(th);
  free(args);
  if (examined) *examined = total;
  return dt > 0 ? total / dt : 0;
}
