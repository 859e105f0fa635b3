/*
 * CPU oracle for the onion-gen candidate-key search.  TEST INFRASTRUCTURE ONLY:
 * the parity tests, the smoke check and the CPU-baseline leg of the benchmark
 * link it as the checker.  Nothing in the product library calls it.
 *
 * Each routine restates the reference's Rust engine (paths relative to the
 * reference repository root):
 *   field   : src/field.rs        (radix 2^64 x 4, folded with 2^256 = 38)
 *   curve   : src/curve.rs        (extended coordinates, cached addend)
 *   batch   : src/batch.rs        (BatchEngine: chain + one inversion per batch)
 *   pairs   : src/pairs.rs        (PairEngine: two candidates per affine offset)
 *   key     : src/key.rs          (seed -> SHA-512 -> clamp, pack, checksum, address)
 *   base32  : src/base32.rs
 *   filter  : src/filter.rs       (parse, naive matcher, bitmap index, length buckets)
 *   run     : src/run.rs:211      (block_seed)
 * SHA-512, SHA3-256 and ChaCha20 come from OpenSSL's libcrypto (textbook library
 * routines, FIPS 180-4 / FIPS 202 / RFC 8439).
 */
#ifndef OG_ORACLE_H
#define OG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field: src/field.rs ---------------------------------------------- */
typedef struct { uint64_t v[4]; } ofe;

void ofe_from_bytes(ofe *r, const uint8_t b[32]);
void ofe_to_bytes(uint8_t out[32], const ofe *a);
void ofe_add(ofe *r, const ofe *a, const ofe *b);
void ofe_sub(ofe *r, const ofe *a, const ofe *b);
void ofe_mul(ofe *r, const ofe *a, const ofe *b);
void ofe_sq(ofe *r, const ofe *a);
void ofe_invert(ofe *r, const ofe *a);
int ofe_is_negative(const ofe *a);

/* byte-level wrappers used by the Python tests */
void oracle_fe_mul(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);
void oracle_fe_add(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);
void oracle_fe_sub(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);
void oracle_fe_invert(const uint8_t *a, uint8_t *out, size_t n);
void oracle_fe_canon(const uint8_t *a, uint8_t *out, size_t n);

/* ---- curve: src/curve.rs ---------------------------------------------- */
typedef struct { ofe x, y, z, t; } opoint;
typedef struct { ofe ypx, ymx, z, t2d; } ocached;

void opoint_identity(opoint *p);
void opoint_basepoint(opoint *p);
void ocurve_two_d(ofe *r);
void ocurve_eight_cached(ocached *c);
void ocurve_add(opoint *r, const opoint *p, const ocached *q); /* add + to_p3 */
void ocurve_dbl(opoint *r, const opoint *p);                   /* dbl + to_p3 */
void ocurve_to_cached(ocached *c, const opoint *p);
void ocurve_scalar_base_mult(opoint *r, const uint8_t scalar[32]);
void ocurve_pack(uint8_t out[32], const opoint *p); /* key::pack */
/* 128 bytes: X, Y, Z, T as canonical little-endian field encodings */
void opoint_to_bytes(uint8_t out[128], const opoint *p);
void opoint_from_bytes(opoint *p, const uint8_t in[128]);

/* ---- key: src/key.rs; base32: src/base32.rs; run: src/run.rs --------- */
void okey_secret_scalar(const uint8_t seed[32], uint8_t sk[32]);
void okey_expanded_secret(const uint8_t seed[32], uint8_t out[64]);
void okey_public_key(const uint8_t seed[32], uint8_t pk[32]);
void okey_checksum(const uint8_t pk[32], uint8_t out[2]);
void okey_address_bytes(const uint8_t pk[32], uint8_t out[35]);
void okey_address(const uint8_t pk[32], char out[57]); /* NUL-terminated */
void okey_add_to_scalar(uint8_t sk[32], uint64_t delta);
int okey_clamping_intact(const uint8_t sk[32]);
int okey_expanded_secret_at_offset(const uint8_t seed[32], uint64_t offset, uint8_t out[64]);
void okey_start_point(const uint8_t seed[32], uint8_t xyzt[128]);
void obase32_encode(const uint8_t *bytes, size_t n, char *out); /* NUL-terminated */
void orun_block_seed(const uint8_t root[32], uint64_t index, uint8_t out[32]);
/* The harness convention for "ChaCha20 seeded n": rand_core's seed_from_u64
 * (PCG32 expansion of n into a 32-byte key) then the ChaCha20 keystream with
 * nonce 0, counter 0.  Writes `len` keystream bytes. */
void ochacha_stream(uint64_t seed, uint8_t *out, size_t len);
void osha3_256(const uint8_t *in, size_t n, uint8_t out[32]);
void osha512(const uint8_t *in, size_t n, uint8_t out[64]);

/* ---- batch engine: src/batch.rs --------------------------------------- */
typedef struct obatch obatch;
obatch *obatch_new(size_t width);
void obatch_free(obatch *e);
size_t obatch_width(const obatch *e);
void obatch_run(obatch *e, opoint *acc);
const uint8_t *obatch_packed(const obatch *e, size_t i); /* sign bit clear */
void obatch_packed_with_sign(const obatch *e, size_t i, uint8_t out[32]);

/* ---- pair engine: src/pairs.rs (the reference's default CPU engine) --- */
struct ofilterset;
typedef struct opairs opairs;
/* batch: the request of PairEngine::new; a round yields 2*(batch/2)+1 candidates */
opairs *opairs_new(size_t batch);
void opairs_free(opairs *e);
size_t opairs_batch_size(const opairs *e);
/* PairEngine::begin: moves the block's base point to the first centre. */
void opairs_begin(opairs *e, opoint *acc);
/* PairEngine::run with the three-form limb probe of the set's bitmap (every
 * candidate survives when probe is NULL); returns the survivor count. */
size_t opairs_run(opairs *e, opoint *acc, const struct ofilterset *probe);
/* Survivor i of the last round: its index in the round and canonical y. */
void opairs_survivor(const opairs *e, size_t i, uint32_t *index, uint8_t packed[32]);
/* PairEngine::delta: the scalar delta of a round index. */
uint64_t opairs_delta(const opairs *e, uint32_t index);

/* ---- filters: src/filter.rs ------------------------------------------- */
typedef struct ofilterset ofilterset;
/* Parses newline-separated filters ("abc", "prefix:abc", "suffix:xyd",
 * "contains:abcde", with '?' and '[..]' classes).  Returns NULL and writes a
 * message on a rejected filter. */
ofilterset *ofilter_parse(const char *text, char *err, size_t errcap);
void ofilter_free(ofilterset *s);
size_t ofilter_count(const ofilterset *s);           /* after dedupe/absorb */
const char *ofilter_canonical(const ofilterset *s, size_t i);
/* Prefix bitmap (FilterSet::index_words): returns bits, 0 if none; words may be NULL to query size. */
uint32_t ofilter_index(const ofilterset *s, uint32_t want_bits, uint64_t *words, size_t nwords);
/* Naive matcher over the whole 35-byte address: the oracle's definition of a hit.
 * Returns the index of the first matching filter (sorted order) or -1. */
int ofilter_match_address(const ofilterset *s, const uint8_t addr[35]);
/* Key-side matcher for literal prefixes through the bitmap and length buckets
 * (FilterSet::match_key restricted to literal prefixes); -1 if none. */
int ofilter_match_key(const ofilterset *s, const uint8_t packed[32]);
int ofilter_needs_checksum(const ofilterset *s);
/* The bitmap test of a key's leading 64 bits (big-endian); 1 without an index. */
int ofilter_may_match_head(const ofilterset *s, uint64_t head);

/* ---- the search oracle ------------------------------------------------ */
typedef struct {
  uint64_t k;          /* candidate index: scalar delta 8k from the start scalar */
  uint8_t pk[32];      /* packed public key, sign bit included */
  char address[57];    /* 56 base32 symbols + NUL */
  int32_t filter;      /* index of the first matching filter */
} ohit;

/* Walks k = k0 .. k0+n-1 from `start_xyzt` (the point of k = 0) with a BatchEngine
 * of `width`, and records every candidate whose address matches the set.
 * Returns the number of hits (may exceed cap; only cap are written). */
size_t osearch(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
               const ofilterset *s, ohit *hits, size_t cap);
/* osearch split over `threads` threads, each on its own contiguous part of the
 * range; the same hits in the same order. */
size_t osearch_mt(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
                  const ofilterset *s, int threads, ohit *hits, size_t cap);
/* XOR and 64-bit wrapping sum of the first 8 bytes of every packed y (sign
 * clear) over k0 .. k0+n-1: a digest over all candidates, hits or not. */
void odigest(const uint8_t start_xyzt[128], uint64_t k0, uint64_t n, size_t width,
             uint64_t out[2]);
/* CPU throughput harness: `threads` independent chains of BatchEngine(width)
 * for `seconds`, filter `s` (literal prefixes, bitmap path).  Returns
 * candidates/s; *examined receives the count. */
double obench_batch(const uint8_t seed[32], size_t width, int threads, double seconds,
                    const ofilterset *s, uint64_t *examined);
/* The same harness over PairEngine(batch) with the limb probe. */
double obench_pairs(const uint8_t seed[32], size_t batch, int threads, double seconds,
                    const ofilterset *s, uint64_t *examined);

#ifdef __cplusplus
}
#endif
#endif
