/*
 * CPU oracle, arithmetic half.  TEST INFRASTRUCTURE ONLY (see og_oracle.h).
 *
 * Field:  src/field.rs:23-255   radix 2^64 x 4, values kept below 2^256 and folded
 *                               with 2^256 = 38 (mod p); canonical form by three
 *                               trial subtractions of p (src/field.rs:46-69).
 * Curve:  src/curve.rs:28-198   ed25519 extended coordinates.
 * Batch:  src/batch.rs:44-140   BatchEngine.
 * Key:    src/key.rs:23-206     seed -> SHA-512 -> clamp; pack; checksum; address.
 * Base32: src/base32.rs:8-88.
 * Run:    src/run.rs:211-220    block_seed.
 */
#include "og_oracle.h"

#include <openssl/evp.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static const uint64_t P[4] = {0xffffffffffffffedull, 0xffffffffffffffffull,
                              0xffffffffffffffffull, 0x7fffffffffffffffull};

/* ------------------------------------------------------------------ field */

void ofe_from_bytes(ofe *r, const uint8_t b[32]) {
  /* src/field.rs:36-44: little-endian limbs, top bit dropped */
  for (int k = 0; k < 4; k++) {
    uint64_t w = 0;
    for (int i = 7; i >= 0; i--) w = (w << 8) | b[8 * k + i];
    r->v[k] = w;
  }
  r->v[3] &= 0x7fffffffffffffffull;
}

void ofe_to_bytes(uint8_t out[32], const ofe *a) {
  /* src/field.rs:46-69: three conditional subtractions of p */
  uint64_t x[4] = {a->v[0], a->v[1], a->v[2], a->v[3]};
  for (int round = 0; round < 3; round++) {
    uint64_t t[4], borrow = 0;
    for (int k = 0; k < 4; k++) {
      u128 d = (u128)x[k] - P[k] - borrow;
      t[k] = (uint64_t)d;
      borrow = (uint64_t)(d >> 127); /* 1 when the subtraction wrapped */
    }
    if (!borrow) memcpy(x, t, sizeof x);
  }
  for (int k = 0; k < 4; k++)
    for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(x[k] >> (8 * i));
}

/* src/field.rs:271-290 add_small: fold a carry out of the top back in as 38 */
static void add_small(uint64_t r[4], uint64_t v) {
  uint64_t carry = v;
  for (int pass = 0; pass < 2 && carry; pass++) {
    uint64_t c = carry;
    for (int k = 0; k < 4 && c; k++) {
      uint64_t s = r[k] + c;
      c = s < r[k];
      r[k] = s;
    }
    carry = 38 * c;
  }
}

/* src/field.rs:292-310 sub_small */
static void sub_small(uint64_t r[4], uint64_t v) {
  uint64_t borrow = v;
  for (int pass = 0; pass < 2 && borrow; pass++) {
    uint64_t b = borrow;
    for (int k = 0; k < 4 && b; k++) {
      uint64_t d = r[k] - b;
      b = r[k] < b;
      r[k] = d;
    }
    borrow = 38 * b;
  }
}

void ofe_add(ofe *r, const ofe *a, const ofe *b) {
  /* src/field.rs:130-144 */
  uint64_t out[4];
  u128 c = 0;
  for (int k = 0; k < 4; k++) {
    c += (u128)a->v[k] + b->v[k];
    out[k] = (uint64_t)c;
    c >>= 64;
  }
  add_small(out, 38 * (uint64_t)c);
  memcpy(r->v, out, sizeof out);
}

void ofe_sub(ofe *r, const ofe *a, const ofe *b) {
  /* src/field.rs:147-160 */
  uint64_t out[4], borrow = 0;
  for (int k = 0; k < 4; k++) {
    u128 d = (u128)a->v[k] - b->v[k] - borrow;
    out[k] = (uint64_t)d;
    borrow = (uint64_t)(d >> 127);
  }
  sub_small(out, 38 * borrow);
  memcpy(r->v, out, sizeof out);
}

void ofe_mul(ofe *r, const ofe *a, const ofe *b) {
  /* src/field.rs:82-117: 4x4 schoolbook, one carry chain per row (mac, :234) */
  uint64_t w[8] = {0};
  for (int i = 0; i < 4; i++) {
    uint64_t carry = 0;
    for (int j = 0; j < 4; j++) {
      u128 t = (u128)a->v[i] * b->v[j] + w[i + j] + carry;
      w[i + j] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
    w[i + 4] = carry;
  }
  /* fold (src/field.rs:245-268): low + 38*high, then two unconditional passes */
  uint64_t out[4];
  u128 c = 0;
  for (int k = 0; k < 4; k++) {
    u128 t = (u128)w[k] + (u128)38 * w[k + 4] + c;
    out[k] = (uint64_t)t;
    c = t >> 64;
  }
  uint64_t v = 38 * (uint64_t)c;
  for (int pass = 0; pass < 2; pass++) {
    for (int k = 0; k < 4; k++) {
      uint64_t s = out[k] + v;
      v = s < out[k];
      out[k] = s;
    }
    v *= 38;
  }
  memcpy(r->v, out, sizeof out);
}

void ofe_sq(ofe *r, const ofe *a) { ofe_mul(r, a, a); }

static void sq_n(ofe *r, const ofe *a, int n) {
  *r = *a;
  for (int i = 0; i < n; i++) ofe_sq(r, r);
}

void ofe_invert(ofe *r, const ofe *z1) {
  /* src/field.rs:163-220: x^(p-2) by the ref10 chain, 254 sq + 11 mul */
  ofe z2, z8, z9, z11, z22, z5, z10, z20, z40, z50, z100, z200, t;
  ofe_sq(&z2, z1);
  sq_n(&z8, &z2, 2);
  ofe_mul(&z9, z1, &z8);
  ofe_mul(&z11, &z2, &z9);
  ofe_sq(&z22, &z11);
  ofe_mul(&z5, &z9, &z22);
  sq_n(&t, &z5, 5);
  ofe_mul(&z10, &t, &z5);
  sq_n(&t, &z10, 10);
  ofe_mul(&z20, &t, &z10);
  sq_n(&t, &z20, 20);
  ofe_mul(&z40, &t, &z20);
  sq_n(&t, &z40, 10);
  ofe_mul(&z50, &t, &z10);
  sq_n(&t, &z50, 50);
  ofe_mul(&z100, &t, &z50);
  sq_n(&t, &z100, 100);
  ofe_mul(&z200, &t, &z100);
  sq_n(&t, &z200, 50);
  ofe_mul(&t, &t, &z50);
  sq_n(&t, &t, 5);
  ofe_mul(r, &t, &z11);
}

int ofe_is_negative(const ofe *a) {
  uint8_t b[32];
  ofe_to_bytes(b, a);
  return b[0] & 1;
}

static int ofe_equal(const ofe *a, const ofe *b) {
  uint8_t x[32], y[32];
  ofe_to_bytes(x, a);
  ofe_to_bytes(y, b);
  return memcmp(x, y, 32) == 0;
}

/* Raw limbs in, raw limbs out: the byte wrappers take 32-byte little-endian
 * values WITHOUT dropping the top bit, so unreduced inputs in [p, 2^256) can be
 * exercised; outputs are canonical encodings. */
static void raw_load(ofe *r, const uint8_t *b) {
  for (int k = 0; k < 4; k++) {
    uint64_t w = 0;
    for (int i = 7; i >= 0; i--) w = (w << 8) | b[8 * k + i];
    r->v[k] = w;
  }
}

void oracle_fe_mul(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    ofe x, y, z;
    raw_load(&x, a + 32 * i);
    raw_load(&y, b + 32 * i);
    ofe_mul(&z, &x, &y);
    ofe_to_bytes(out + 32 * i, &z);
  }
}
void oracle_fe_add(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    ofe x, y, z;
    raw_load(&x, a + 32 * i);
    raw_load(&y, b + 32 * i);
    ofe_add(&z, &x, &y);
    ofe_to_bytes(out + 32 * i, &z);
  }
}
void oracle_fe_sub(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    ofe x, y, z;
    raw_load(&x, a + 32 * i);
    raw_load(&y, b + 32 * i);
    ofe_sub(&z, &x, &y);
    ofe_to_bytes(out + 32 * i, &z);
  }
}
void oracle_fe_invert(const uint8_t *a, uint8_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    ofe x, z;
    raw_load(&x, a + 32 * i);
    ofe_invert(&z, &x);
    ofe_to_bytes(out + 32 * i, &z);
  }
}
void oracle_fe_canon(const uint8_t *a, uint8_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    ofe x;
    raw_load(&x, a + 32 * i);
    ofe_to_bytes(out + 32 * i, &x);
  }
}

/* ------------------------------------------------------------------ curve */

static void hex32(uint8_t out[32], const char *s) {
  for (int i = 0; i < 32; i++) {
    int hi = s[2 * i], lo = s[2 * i + 1];
    hi = hi <= '9' ? hi - '0' : hi - 'a' + 10;
    lo = lo <= '9' ? lo - '0' : lo - 'a' + 10;
    out[i] = (uint8_t)(hi * 16 + lo);
  }
}

static void fe_hex(ofe *r, const char *s) {
  uint8_t b[32];
  hex32(b, s);
  ofe_from_bytes(r, b);
}

/* src/curve.rs:28-39 */
#define BASE_X "1ad5258f602d56c9b2a7259560c72c695cdcd6fd31e2a4c0fe536ecdd3366921"
#define BASE_Y "5866666666666666666666666666666666666666666666666666666666666666"
#define EIGHT_YPX "8f3edd046659b7592c7088e27703b36c23c3d95e669c33b12fe5bc6160e71509"
#define EIGHT_YMX "d93492f3ed5da7e2f958b5e180763d96fb233c6eac41272cc3010e32a124903a"
#define EIGHT_T2D "1a91a2c9d9f5c1e7d7a7cc8b7871a3b8322ab60e19126463954ecc2e5c7c9026"
#define TWO_D "59f1b226949bd6eb56b183829a14e00030d1f3eef2808e19e7fcdf56dcd90624"

void opoint_identity(opoint *p) {
  memset(p, 0, sizeof *p);
  p->y.v[0] = 1;
  p->z.v[0] = 1;
}

void opoint_basepoint(opoint *p) {
  fe_hex(&p->x, BASE_X);
  fe_hex(&p->y, BASE_Y);
  memset(&p->z, 0, sizeof p->z);
  p->z.v[0] = 1;
  ofe_mul(&p->t, &p->x, &p->y);
}

void ocurve_two_d(ofe *r) { fe_hex(r, TWO_D); }

void ocurve_eight_cached(ocached *c) {
  fe_hex(&c->ypx, EIGHT_YPX);
  fe_hex(&c->ymx, EIGHT_YMX);
  memset(&c->z, 0, sizeof c->z);
  c->z.v[0] = 1;
  fe_hex(&c->t2d, EIGHT_T2D);
}

/* src/curve.rs:168-175 to_p3 of a P1P1 given as (x, y, z, t) */
static void to_p3(opoint *r, const ofe *x, const ofe *y, const ofe *z, const ofe *t) {
  ofe X, Y, Z, T;
  ofe_mul(&X, x, t);
  ofe_mul(&Y, y, z);
  ofe_mul(&Z, z, t);
  ofe_mul(&T, x, y);
  r->x = X;
  r->y = Y;
  r->z = Z;
  r->t = T;
}

void ocurve_add(opoint *r, const opoint *p, const ocached *q) {
  /* src/curve.rs:130-145 add, then to_p3 */
  ofe a, b, zz, yy, tt, t0, X, Y, Z, T;
  ofe_add(&a, &p->y, &p->x);
  ofe_sub(&b, &p->y, &p->x);
  ofe_mul(&zz, &a, &q->ypx);
  ofe_mul(&yy, &b, &q->ymx);
  ofe_mul(&tt, &q->t2d, &p->t);
  ofe_mul(&t0, &p->z, &q->z);
  ofe_add(&t0, &t0, &t0);
  ofe_sub(&X, &zz, &yy);
  ofe_add(&Y, &zz, &yy);
  ofe_add(&Z, &t0, &tt);
  ofe_sub(&T, &t0, &tt);
  to_p3(r, &X, &Y, &Z, &T);
}

void ocurve_dbl(opoint *r, const opoint *p) {
  /* src/curve.rs:148-164 dbl, then to_p3 */
  ofe xx, zz, tt, yy, t0, y, z, X, T;
  ofe_sq(&xx, &p->x);
  ofe_sq(&zz, &p->y);
  ofe_sq(&tt, &p->z);
  ofe_add(&tt, &tt, &tt);
  ofe_add(&yy, &p->x, &p->y);
  ofe_sq(&t0, &yy);
  ofe_add(&y, &zz, &xx);
  ofe_sub(&z, &zz, &xx);
  ofe_sub(&X, &t0, &y);
  ofe_sub(&T, &tt, &z);
  to_p3(r, &X, &y, &z, &T);
}

void ocurve_to_cached(ocached *c, const opoint *p) {
  /* src/curve.rs:177-184 */
  ofe two_d;
  ocurve_two_d(&two_d);
  ofe_add(&c->ypx, &p->y, &p->x);
  ofe_sub(&c->ymx, &p->y, &p->x);
  c->z = p->z;
  ofe_mul(&c->t2d, &p->t, &two_d);
}

void ocurve_scalar_base_mult(opoint *r, const uint8_t scalar[32]) {
  /* src/curve.rs:188-198: MSB-first double-and-add over all 256 bits */
  opoint base, acc;
  ocached bc;
  opoint_basepoint(&base);
  ocurve_to_cached(&bc, &base);
  opoint_identity(&acc);
  for (int i = 255; i >= 0; i--) {
    ocurve_dbl(&acc, &acc);
    if ((scalar[i / 8] >> (i % 8)) & 1) ocurve_add(&acc, &acc, &bc);
  }
  *r = acc;
}

void ocurve_pack(uint8_t out[32], const opoint *p) {
  /* src/key.rs:55-62 */
  ofe zinv, x, y;
  ofe_invert(&zinv, &p->z);
  ofe_mul(&y, &p->y, &zinv);
  ofe_mul(&x, &p->x, &zinv);
  ofe_to_bytes(out, &y);
  out[31] ^= (uint8_t)(ofe_is_negative(&x) << 7);
}

void opoint_to_bytes(uint8_t out[128], const opoint *p) {
  ofe_to_bytes(out, &p->x);
  ofe_to_bytes(out + 32, &p->y);
  ofe_to_bytes(out + 64, &p->z);
  ofe_to_bytes(out + 96, &p->t);
}

void opoint_from_bytes(opoint *p, const uint8_t in[128]) {
  raw_load(&p->x, in);
  raw_load(&p->y, in + 32);
  raw_load(&p->z, in + 64);
  raw_load(&p->t, in + 96);
}

/* ------------------------------------------------------------ hashes, key */

void osha512(const uint8_t *in, size_t n, uint8_t out[64]) {
  unsigned len = 64;
  EVP_Digest(in, n, out, &len, EVP_sha512(), NULL);
}

void osha3_256(const uint8_t *in, size_t n, uint8_t out[32]) {
  unsigned len = 32;
  EVP_Digest(in, n, out, &len, EVP_sha3_256(), NULL);
}

static void clamp(uint8_t sk[32]) {
  /* src/key.rs:47-51 */
  sk[0] &= 248;
  sk[31] &= 63;
  sk[31] |= 64;
}

void okey_secret_scalar(const uint8_t seed[32], uint8_t sk[32]) {
  uint8_t d[64];
  osha512(seed, 32, d);
  memcpy(sk, d, 32);
  clamp(sk);
}

void okey_expanded_secret(const uint8_t seed[32], uint8_t out[64]) {
  osha512(seed, 32, out);
  clamp(out);
}

void okey_public_key(const uint8_t seed[32], uint8_t pk[32]) {
  uint8_t sk[32];
  opoint a;
  okey_secret_scalar(seed, sk);
  ocurve_scalar_base_mult(&a, sk);
  ocurve_pack(pk, &a);
}

void okey_start_point(const uint8_t seed[32], uint8_t xyzt[128]) {
  uint8_t sk[32];
  opoint a;
  okey_secret_scalar(seed, sk);
  ocurve_scalar_base_mult(&a, sk);
  opoint_to_bytes(xyzt, &a);
}

void okey_checksum(const uint8_t pk[32], uint8_t out[2]) {
  /* src/key.rs:106-113: SHA3-256(".onion checksum" || pk || 0x03)[0..2] */
  uint8_t buf[48], d[32];
  memcpy(buf, ".onion checksum", 15);
  memcpy(buf + 15, pk, 32);
  buf[47] = 0x03;
  osha3_256(buf, 48, d);
  out[0] = d[0];
  out[1] = d[1];
}

void okey_address_bytes(const uint8_t pk[32], uint8_t out[35]) {
  memcpy(out, pk, 32);
  okey_checksum(pk, out + 32);
  out[34] = 0x03;
}

static const char ALPHA[] = "abcdefghijklmnopqrstuvwxyz234567";

void obase32_encode(const uint8_t *bytes, size_t n, char *out) {
  /* src/base32.rs:42-62: MSB-first 5-bit groups, zero-padded tail */
  unsigned acc = 0, bits = 0;
  size_t o = 0;
  for (size_t i = 0; i < n; i++) {
    acc = (acc << 8) | bytes[i];
    bits += 8;
    while (bits >= 5) {
      bits -= 5;
      out[o++] = ALPHA[(acc >> bits) & 31];
    }
  }
  if (bits) out[o++] = ALPHA[(acc << (5 - bits)) & 31];
  out[o] = 0;
}

void okey_address(const uint8_t pk[32], char out[57]) {
  uint8_t ab[35];
  okey_address_bytes(pk, ab);
  obase32_encode(ab, 35, out);
}

void okey_add_to_scalar(uint8_t sk[32], uint64_t delta) {
  /* src/key.rs:185-196 */
  u128 carry = delta;
  for (int i = 0; i < 32 && carry; i++) {
    carry += sk[i];
    sk[i] = (uint8_t)carry;
    carry >>= 8;
  }
}

int okey_clamping_intact(const uint8_t sk[32]) {
  /* src/key.rs:204-206 */
  return (sk[0] & 248) == sk[0] && (((sk[31] & 63) | 64) == sk[31]);
}

int okey_expanded_secret_at_offset(const uint8_t seed[32], uint64_t offset, uint8_t out[64]) {
  /* src/key.rs:167-177 */
  okey_expanded_secret(seed, out);
  okey_add_to_scalar(out, offset);
  return okey_clamping_intact(out);
}

void orun_block_seed(const uint8_t root[32], uint64_t index, uint8_t out[32]) {
  /* src/run.rs:211-220 */
  uint8_t buf[18 + 32 + 8], d[64];
  memcpy(buf, "onion-gen block v1", 18);
  memcpy(buf + 18, root, 32);
  for (int i = 0; i < 8; i++) buf[50 + i] = (uint8_t)(index >> (8 * i));
  osha512(buf, sizeof buf, d);
  memcpy(out, d, 32);
}

void ochacha_stream(uint64_t seed, uint8_t *out, size_t len) {
  /* rand_core 0.6 SeedableRng::seed_from_u64: PCG32 XSH-RR words, little-endian */
  uint8_t key[32];
  uint64_t state = seed;
  for (int w = 0; w < 8; w++) {
    state = state * 6364136223846793005ull + 11634580027462260723ull;
    uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    uint32_t x = (xs >> rot) | (xs << ((32 - rot) & 31));
    for (int b = 0; b < 4; b++) key[4 * w + b] = (uint8_t)(x >> (8 * b));
  }
  uint8_t iv[16] = {0}; /* counter 0, nonce 0 */
  EVP_CIPHER_CTX *ctx = EVP_CIPHER_CTX_new();
  EVP_EncryptInit_ex(ctx, EVP_chacha20(), NULL, key, iv);
  uint8_t *zero = calloc(len ? len : 1, 1);
  int outl = 0;
  EVP_EncryptUpdate(ctx, out, &outl, zero, (int)len);
  free(zero);
  EVP_CIPHER_CTX_free(ctx);
}

/* ------------------------------------------------------------------ batch */

struct obatch {
  size_t n;
  ofe *x, *y, *z, *scratch;
  uint8_t (*packed)[32];
  ocached step;
};

obatch *obatch_new(size_t width) {
  obatch *e = calloc(1, sizeof *e);
  e->n = width;
  e->x = calloc(width, sizeof(ofe));
  e->y = calloc(width, sizeof(ofe));
  e->z = calloc(width, sizeof(ofe));
  e->scratch = calloc(width, sizeof(ofe));
  e->packed = calloc(width, 32);
  ocurve_eight_cached(&e->step);
  return e;
}

void obatch_free(obatch *e) {
  if (!e) return;
  free(e->x);
  free(e->y);
  free(e->z);
  free(e->scratch);
  free(e->packed);
  free(e);
}

size_t obatch_width(const obatch *e) { return e->n; }

void obatch_run(obatch *e, opoint *acc) {
  /* src/batch.rs:78-91: store (X, Y, Z), then acc += 8G */
  for (size_t i = 0; i < e->n; i++) {
    e->x[i] = acc->x;
    e->y[i] = acc->y;
    e->z[i] = acc->z;
    ocurve_add(acc, acc, &e->step);
  }
  /* src/batch.rs:102-122 pack: forward prefix products, one inversion, backward */
  ofe a;
  memset(&a, 0, sizeof a);
  a.v[0] = 1;
  for (size_t i = 0; i < e->n; i++) {
    e->scratch[i] = a;
    ofe_mul(&a, &a, &e->z[i]);
  }
  ofe_invert(&a, &a);
  for (size_t i = e->n; i-- > 0;) {
    ofe next, zinv, y;
    ofe_mul(&next, &a, &e->z[i]);
    ofe_mul(&zinv, &a, &e->scratch[i]);
    a = next;
    e->z[i] = zinv;
    ofe_mul(&y, &e->y[i], &zinv);
    ofe_to_bytes(e->packed[i], &y);
  }
}

const uint8_t *obatch_packed(const obatch *e, size_t i) { return e->packed[i]; }

void obatch_packed_with_sign(const obatch *e, size_t i, uint8_t out[32]) {
  /* src/batch.rs:129-139 */
  ofe x;
  memcpy(out, e->packed[i], 32);
  ofe_mul(&x, &e->x[i], &e->z[i]);
  out[31] ^= (uint8_t)(ofe_is_negative(&x) << 7);
}

/* internal helper for tests that want the equality by value */
int oracle_fe_equal_bytes(const uint8_t *a, const uint8_t *b) {
  ofe x, y;
  raw_load(&x, a);
  raw_load(&y, b);
  return ofe_equal(&x, &y);
}
