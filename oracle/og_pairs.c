/*
 * PairEngine restated (src/pairs.rs:38-282), the reference's default CPU
 * engine.  TEST INFRASTRUCTURE ONLY: the CPU baseline of the benchmark and
 * the tests that pin it to the BatchEngine oracle.
 *
 * A round reads two candidates out of one affine centre P and each table
 * offset Q(m) = 8m*G with the dedicated Edwards formula (src/pairs.rs:1-19):
 *     y(P+Q) = (x1y1 - x2y2) / (x1y2 - y1x2),  y(P-Q) = (x1y1 + x2y2) / (x1y2 + y1x2)
 * numerators folded with the running product of the denominators, one
 * inversion per round, the centre itself as y/1 (src/pairs.rs:171-230).
 * Round r covers deltas centre - 8*half .. centre + 8*half; centres advance by
 * 8*(2*half+1) from 8*(half+1), so the deltas tile 8, 16, 24, ...
 */
#include <stdlib.h>
#include <string.h>

#include "og_oracle.h"

typedef struct {
  ofe x, y, xy;
} oaffine;

struct opairs {
  size_t half, n;
  oaffine *table; /* table[j] = 8(j+1)*G */
  ocached stride, to_first;
  ofe *num, *den;
  oaffine centre;
  uint64_t centre_delta;
  uint32_t *surv_index;
  uint8_t *surv_packed;
  size_t nsurv;
};

static void fe_one(ofe *r) {
  memset(r, 0, sizeof *r);
  r->v[0] = 1;
}

static void to_affine(oaffine *a, const opoint *p) {
  /* src/pairs.rs:302-311 */
  ofe zinv;
  ofe_invert(&zinv, &p->z);
  ofe_mul(&a->x, &p->x, &zinv);
  ofe_mul(&a->y, &p->y, &zinv);
  ofe_mul(&a->xy, &a->x, &a->y);
}

static void multiple_of_eight_g(opoint *r, uint64_t k) {
  /* src/pairs.rs:293-299: (8k) as a 128-bit little-endian scalar */
  uint8_t scalar[32] = {0};
  unsigned __int128 product = (unsigned __int128)k * 8;
  for (int i = 0; i < 16; i++) scalar[i] = (uint8_t)(product >> (8 * i));
  ocurve_scalar_base_mult(r, scalar);
}

opairs *opairs_new(size_t batch) {
  if (batch < 2) return NULL;
  opairs *e = calloc(1, sizeof *e);
  e->half = batch / 2;
  e->n = 2 * e->half + 1;
  e->table = calloc(e->half, sizeof *e->table);
  e->num = calloc(e->n, sizeof *e->num);
  e->den = calloc(e->n, sizeof *e->den);
  e->surv_index = calloc(e->n, sizeof *e->surv_index);
  e->surv_packed = calloc(e->n, 32);
  /* Q(1)..Q(half) by repeated addition, then one inversion for all of them
   * (src/pairs.rs:93-103, 313-343) */
  opoint *pts = calloc(e->half, sizeof *pts);
  ocached step;
  ocurve_eight_cached(&step);
  opoint p;
  opoint_identity(&p);
  ocurve_add(&p, &p, &step);
  for (size_t j = 0; j < e->half; j++) {
    pts[j] = p;
    ocurve_add(&p, &p, &step);
  }
  ofe *prefix = calloc(e->half, sizeof *prefix);
  ofe running;
  fe_one(&running);
  for (size_t i = 0; i < e->half; i++) {
    prefix[i] = running;
    ofe_mul(&running, &running, &pts[i].z);
  }
  ofe inv;
  ofe_invert(&inv, &running);
  for (size_t i = e->half; i-- > 0;) {
    ofe next, zinv;
    ofe_mul(&next, &inv, &pts[i].z);
    ofe_mul(&zinv, &inv, &prefix[i]);
    inv = next;
    ofe_mul(&e->table[i].x, &pts[i].x, &zinv);
    ofe_mul(&e->table[i].y, &pts[i].y, &zinv);
    ofe_mul(&e->table[i].xy, &e->table[i].x, &e->table[i].y);
  }
  free(prefix);
  free(pts);
  opoint m;
  multiple_of_eight_g(&m, 2 * e->half + 1);
  ocurve_to_cached(&e->stride, &m);
  multiple_of_eight_g(&m, e->half + 1);
  ocurve_to_cached(&e->to_first, &m);
  return e;
}

void opairs_free(opairs *e) {
  if (!e) return;
  free(e->table);
  free(e->num);
  free(e->den);
  free(e->surv_index);
  free(e->surv_packed);
  free(e);
}

size_t opairs_batch_size(const opairs *e) { return e->n; }

void opairs_begin(opairs *e, opoint *acc) {
  /* src/pairs.rs:137-143 */
  ocurve_add(acc, acc, &e->to_first);
  e->centre_delta = 8 * ((uint64_t)e->half + 1);
}

/* LimbProbe::may_match (src/filter.rs:2619-2631): the limb, plus 19, plus 38 */
static int limb_may_match(const struct ofilterset *s, uint64_t limb0) {
  uint64_t v = limb0;
  for (int i = 0; i < 3; i++) {
    if (ofilter_may_match_head(s, __builtin_bswap64(v))) return 1;
    v += 19;
  }
  return 0;
}

size_t opairs_run(opairs *e, opoint *acc, const struct ofilterset *probe) {
  /* src/pairs.rs:171-230 */
  const size_t half = e->half, n = e->n;
  to_affine(&e->centre, acc);
  const oaffine *pa = &e->centre;
  ofe running;
  fe_one(&running);
  for (size_t j = 0; j < half; j++) {
    const oaffine *q = &e->table[j];
    ofe x1y2, y1x2, t;
    ofe_mul(&x1y2, &pa->x, &q->y);
    ofe_mul(&y1x2, &pa->y, &q->x);
    ofe d0, d1;
    ofe_sub(&d0, &x1y2, &y1x2);
    ofe_sub(&t, &pa->xy, &q->xy);
    ofe_mul(&e->num[2 * j], &t, &running);
    e->den[2 * j] = d0;
    ofe_mul(&running, &running, &d0);
    ofe_add(&d1, &x1y2, &y1x2);
    ofe_add(&t, &pa->xy, &q->xy);
    ofe_mul(&e->num[2 * j + 1], &t, &running);
    e->den[2 * j + 1] = d1;
    ofe_mul(&running, &running, &d1);
  }
  ofe_mul(&e->num[n - 1], &pa->y, &running);
  fe_one(&e->den[n - 1]);
  ofe inv;
  ofe_invert(&inv, &running);
  e->nsurv = 0;
  for (size_t i = n; i-- > 0;) {
    ofe v;
    ofe_mul(&v, &inv, &e->num[i]);
    ofe_mul(&inv, &inv, &e->den[i]);
    if (probe && !limb_may_match(probe, v.v[0])) continue;
    e->surv_index[e->nsurv] = (uint32_t)i;
    ofe_to_bytes(e->surv_packed + 32 * e->nsurv, &v);
    e->nsurv++;
  }
  ocurve_add(acc, acc, &e->stride);
  e->centre_delta += 8 * (2 * (uint64_t)half + 1);
  return e->nsurv;
}

void opairs_survivor(const opairs *e, size_t i, uint32_t *index, uint8_t packed[32]) {
  *index = e->surv_index[i];
  memcpy(packed, e->surv_packed + 32 * i, 32);
}

uint64_t opairs_delta(const opairs *e, uint32_t index) {
  /* src/pairs.rs:239-252: the centre has already moved on by one stride */
  const uint64_t centre = e->centre_delta - 8 * (2 * (uint64_t)e->half + 1);
  if (index == 2 * e->half) return centre;
  const uint64_t m = 8 * ((uint64_t)index / 2 + 1);
  return (index % 2 == 0) ? centre + m : centre - m;
}
