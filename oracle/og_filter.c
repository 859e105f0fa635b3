/*
 * CPU oracle, matching half.  TEST INFRASTRUCTURE ONLY (see og_oracle.h).
 *
 * Restates src/filter.rs:
 *   Pattern::parse / parse_symbols / check_reachable_tail   :298-351, :414-494, :570-601
 *   Filter::parse (bytes + last-byte mask)                   :649-684
 *   Filter::matches_symbols (the naive matcher)              :744-768
 *   Filter::matches_key                                      :819-832
 *   FilterSet::new (sort by canonical text, drop absorbed)   :1326-1338, absorbs :873-882
 *   BitmapIndex::{build, insert, slot, may_match, occupancy} :2321-2604
 *   narrowest_acceptable / TARGET_FALSE_HITS / widths 16..24 :2248-2299
 *   LengthBucket + match_literals                            :2645-2701, :2194-2217
 * Regular expressions (the `regex:` form) are outside this oracle.
 */
#include "og_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ADDRESS_LEN 56
#define KEY_ONLY_SYMBOLS 49

enum { A_START = 0, A_ANY = 1, A_END = 2 };

typedef struct {
  int anchor;
  int len;
  uint32_t sets[ADDRESS_LEN]; /* bit v set: symbol value v accepted */
  char canonical[ADDRESS_LEN + 1];
  int literal;
  uint8_t bytes[32];
  int nbytes;
  uint8_t mask;
} ofilter;

typedef struct {
  uint64_t mask;
  size_t n;
  uint64_t *keys;
  uint32_t *owners;
} obucket;

struct ofilterset {
  size_t n;
  ofilter *f;
  /* literal fast prefixes: indices into f */
  size_t nfast;
  uint32_t *fast;
  uint32_t index_bits; /* 0: none */
  uint64_t *words;
  size_t nbuckets;
  obucket *buckets;
  int needs_checksum;
};

static int sym_value(int c) {
  c = tolower(c);
  if (c >= 'a' && c <= 'z') return c - 'a';
  if (c >= '2' && c <= '7') return c - '2' + 26;
  return -1;
}

static const char ALPHA[] = "abcdefghijklmnopqrstuvwxyz234567";

static int popcount32(uint32_t x) { return __builtin_popcount(x); }

/* symbol m = bits [5m, 5m+5) of a big-endian bit string (src/filter.rs:259-267) */
static int symbol_at(const uint8_t *bytes, size_t nbytes, int position) {
  int bit = position * 5, byte = bit / 8, shift = bit % 8;
  unsigned hi = (unsigned)bytes[byte] << 8;
  unsigned lo = (size_t)(byte + 1) < nbytes ? bytes[byte + 1] : 0;
  return (int)(((hi | lo) >> (11 - shift)) & 31);
}

static int parse_one(const char *raw, ofilter *out, char *err, size_t errcap) {
  char buf[512];
  size_t n = strlen(raw);
  while (n && isspace((unsigned char)raw[n - 1])) n--;
  while (n && isspace((unsigned char)*raw)) raw++, n--;
  if (n >= sizeof buf) n = sizeof buf - 1;
  memcpy(buf, raw, n);
  buf[n] = 0;
  memset(out, 0, sizeof *out);
  const char *body = buf;
  char *colon = strchr(buf, ':');
  if (colon) {
    *colon = 0;
    char name[64];
    size_t m = 0;
    for (const char *p = buf; *p && m < sizeof name - 1; p++)
      if (!isspace((unsigned char)*p)) name[m++] = (char)tolower((unsigned char)*p);
    name[m] = 0;
    if (!strcmp(name, "prefix")) out->anchor = A_START;
    else if (!strcmp(name, "contains")) out->anchor = A_ANY;
    else if (!strcmp(name, "suffix")) out->anchor = A_END;
    else {
      snprintf(err, errcap, "unknown form \"%s\"", name);
      return -1;
    }
    body = colon + 1;
    while (isspace((unsigned char)*body)) body++;
  } else {
    out->anchor = A_START;
    while (*body == '^') body++;
  }
  /* parse_symbols (src/filter.rs:414-494) */
  int len = 0;
  for (size_t i = 0; body[i];) {
    if (len >= ADDRESS_LEN) {
      snprintf(err, errcap, "the filter is too long");
      return -1;
    }
    char c = body[i];
    if (c == '?') {
      out->sets[len++] = 0xffffffffu;
      i++;
    } else if (c == '[') {
      const char *end = strchr(body + i, ']');
      if (!end) {
        snprintf(err, errcap, "unbalanced [ or ] at position %zu", i);
        return -1;
      }
      size_t start = i + 1, stop = (size_t)(end - body);
      uint32_t set = 0;
      for (size_t j = start; j < stop;) {
        int first = sym_value(body[j]);
        if (first < 0) {
          snprintf(err, errcap, "invalid character at position %zu", j);
          return -1;
        }
        if (j + 2 < stop && body[j + 1] == '-') {
          int last = sym_value(body[j + 2]);
          if (last < 0) {
            snprintf(err, errcap, "invalid character at position %zu", j + 2);
            return -1;
          }
          if (last < first) {
            snprintf(err, errcap, "empty range at position %zu", j);
            return -1;
          }
          for (int v = first; v <= last; v++) set |= 1u << v;
          j += 3;
        } else {
          set |= 1u << first;
          j++;
        }
      }
      if (!set) {
        snprintf(err, errcap, "empty character class at position %zu", i);
        return -1;
      }
      out->sets[len++] = set;
      i = stop + 1;
    } else if (c == ']') {
      snprintf(err, errcap, "unbalanced [ or ] at position %zu", i);
      return -1;
    } else {
      int v = sym_value(c);
      if (v < 0) {
        snprintf(err, errcap, "invalid character '%c' at position %zu", c, i);
        return -1;
      }
      out->sets[len++] = 1u << v;
      i++;
    }
  }
  if (len == 0) {
    snprintf(err, errcap, "the filter is empty");
    return -1;
  }
  int all = 1;
  for (int i = 0; i < len; i++) all &= out->sets[i] == 0xffffffffu;
  if (all) {
    snprintf(err, errcap, "the pattern fixes no symbol");
    return -1;
  }
  /* check_reachable_tail (src/filter.rs:570-601) */
  if (out->anchor == A_END) {
    if (!(out->sets[len - 1] & (1u << sym_value('d')))) {
      snprintf(err, errcap, "no address can end this way: the last symbol must be 'd'");
      return -1;
    }
    if (len >= 2) {
      uint32_t ok = (1u << sym_value('a')) | (1u << sym_value('i')) | (1u << sym_value('q')) |
                    (1u << sym_value('y'));
      if (!(out->sets[len - 2] & ok)) {
        snprintf(err, errcap, "no address can end this way: the penultimate symbol must be one of aiqy");
        return -1;
      }
    }
  }
  out->len = len;
  out->literal = 1;
  for (int i = 0; i < len; i++) {
    int low = __builtin_ctz(out->sets[i]);
    out->canonical[i] = ALPHA[low];
    out->literal &= popcount32(out->sets[i]) == 1;
  }
  out->canonical[len] = 0;
  /* Filter::parse (src/filter.rs:649-684): decode the first <= 49 symbols */
  int head = len < KEY_ONLY_SYMBOLS ? len : KEY_ONLY_SYMBOLS;
  unsigned acc = 0, bits = 0;
  int nb = 0;
  for (int i = 0; i < head; i++) {
    acc = (acc << 5) | (unsigned)__builtin_ctz(out->sets[i]);
    bits += 5;
    if (bits >= 8) {
      bits -= 8;
      out->bytes[nb++] = (uint8_t)(acc >> bits);
    }
  }
  if (bits) out->bytes[nb++] = (uint8_t)(acc << (8 - bits));
  int total = head * 5;
  out->nbytes = (total + 7) / 8;
  int rem = total % 8;
  out->mask = rem == 0 ? 0xff : (uint8_t)(0xff << (8 - rem));
  out->bytes[out->nbytes - 1] &= out->mask;
  return 0;
}

static int is_fast(const ofilter *f) {
  /* src/filter.rs:707-711 */
  return f->anchor == A_START && f->literal && f->len <= KEY_ONLY_SYMBOLS;
}

static int needs_checksum(const ofilter *f) {
  /* any placement reaching past the key-only span (src/filter.rs:393-401) */
  switch (f->anchor) {
    case A_START: return f->len > KEY_ONLY_SYMBOLS;
    case A_END: return 1;
    default: return 1; /* the last placement always ends at symbol 55 */
  }
}

/* Filter::matches_symbols over the 35 address bytes (have_checksum = true) */
static int matches_address(const ofilter *f, const uint8_t addr[35]) {
  int first = 0, last = 0;
  if (f->anchor == A_START) first = last = 0;
  else if (f->anchor == A_END) first = last = ADDRESS_LEN - f->len;
  else {
    first = 0;
    last = ADDRESS_LEN - f->len;
  }
  for (int s = first; s <= last; s++) {
    int ok = 1;
    for (int i = 0; i < f->len && ok; i++) ok = (f->sets[i] >> symbol_at(addr, 35, s + i)) & 1;
    if (ok) return 1;
  }
  return 0;
}

/* Filter::matches_key (src/filter.rs:819-832) */
static int matches_key(const ofilter *f, const uint8_t packed[32]) {
  int last = f->nbytes - 1;
  return memcmp(packed, f->bytes, (size_t)last) == 0 && (packed[last] & f->mask) == f->bytes[last];
}

/* ----------------------------------------------------------- bitmap index */

static size_t slot_of(uint32_t bits, const uint8_t *packed) {
  /* BitmapIndex::slot with offset 0 (src/filter.rs:2538-2549) */
  uint64_t w = 0;
  for (int i = 0; i < 8; i++) w = (w << 8) | packed[i];
  return (size_t)(w >> (64 - bits));
}

static double occupancy(const uint64_t *words, uint32_t bits) {
  size_t nw = ((size_t)1 << bits) / 64 + (bits < 6);
  uint64_t set = 0;
  for (size_t i = 0; i < nw; i++) set += (uint64_t)__builtin_popcountll(words[i]);
  return (double)set / (double)((uint64_t)1 << bits);
}

static uint64_t *build_index(const ofilterset *s, uint32_t bits) {
  /* BitmapIndex::build + insert (src/filter.rs:2334-2350, 2552-2574) */
  if (s->nfast == 0 || bits < 8 || bits > 32) return NULL;
  size_t nw = (((size_t)1 << bits) + 63) / 64;
  uint64_t *words = calloc(nw, 8);
  for (size_t k = 0; k < s->nfast; k++) {
    const ofilter *f = &s->f[s->fast[k]];
    uint8_t probe[32] = {0};
    memcpy(probe, f->bytes, (size_t)f->nbytes);
    uint32_t prefix_bits = (uint32_t)f->len * 5;
    size_t first = slot_of(bits, probe);
    size_t run = prefix_bits >= bits ? 1 : (size_t)1 << (bits - prefix_bits);
    for (size_t sl = first; sl < first + run; sl++) words[sl >> 6] |= 1ull << (sl & 63);
  }
  if (occupancy(words, bits) > 0.25) {
    free(words);
    return NULL;
  }
  return words;
}

typedef struct {
  uint64_t key;
  uint32_t owner;
} kv;

static int cmp_kv(const void *a, const void *b) {
  const kv *x = a, *y = b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (int)x->owner - (int)y->owner;
}

static uint64_t head64(const uint8_t *p) {
  uint64_t w = 0;
  for (int i = 0; i < 8; i++) w = (w << 8) | p[i];
  return w;
}

static void build_buckets(ofilterset *s) {
  /* FilterSet::build_buckets (src/filter.rs:1884-1910), SCAN_LIMIT 64 */
  if (s->nfast <= 64) return;
  for (int sym = 1; sym <= KEY_ONLY_SYMBOLS; sym++) {
    size_t cnt = 0;
    for (size_t k = 0; k < s->nfast; k++) cnt += s->f[s->fast[k]].len == sym;
    if (!cnt) continue;
    uint32_t bits = (uint32_t)(sym * 5 > 64 ? 64 : sym * 5);
    uint64_t mask = bits == 64 ? ~0ull : ~0ull << (64 - bits);
    kv *e = malloc(cnt * sizeof *e);
    size_t j = 0;
    for (size_t k = 0; k < s->nfast; k++) {
      const ofilter *f = &s->f[s->fast[k]];
      if (f->len != sym) continue;
      uint8_t padded[32] = {0};
      memcpy(padded, f->bytes, (size_t)f->nbytes);
      e[j].key = head64(padded) & mask;
      e[j].owner = s->fast[k];
      j++;
    }
    qsort(e, cnt, sizeof *e, cmp_kv);
    s->buckets = realloc(s->buckets, (s->nbuckets + 1) * sizeof *s->buckets);
    obucket *b = &s->buckets[s->nbuckets++];
    b->mask = mask;
    b->n = cnt;
    b->keys = malloc(cnt * 8);
    b->owners = malloc(cnt * 4);
    for (size_t i = 0; i < cnt; i++) {
      b->keys[i] = e[i].key;
      b->owners[i] = e[i].owner;
    }
    free(e);
  }
}

ofilterset *ofilter_parse(const char *text, char *err, size_t errcap) {
  size_t cap = 16, n = 0;
  ofilter *all = malloc(cap * sizeof *all);
  const char *p = text;
  while (*p) {
    const char *e = strchr(p, '\n');
    size_t m = e ? (size_t)(e - p) : strlen(p);
    char line[512];
    if (m >= sizeof line) m = sizeof line - 1;
    memcpy(line, p, m);
    line[m] = 0;
    int blank = 1;
    for (size_t i = 0; i < m; i++) blank &= isspace((unsigned char)line[i]) != 0;
    if (!blank) {
      if (n == cap) all = realloc(all, (cap *= 2) * sizeof *all);
      if (parse_one(line, &all[n], err, errcap) != 0) {
        free(all);
        return NULL;
      }
      n++;
    }
    p = e ? e + 1 : p + m;
  }
  /* FilterSet::new: stable sort by canonical text, then drop what the previous
   * kept filter absorbs (its canonical text is a prefix of this one's). */
  for (size_t i = 1; i < n; i++) { /* insertion sort keeps it stable */
    ofilter t = all[i];
    size_t j = i;
    while (j > 0 && strcmp(all[j - 1].canonical, t.canonical) > 0) {
      all[j] = all[j - 1];
      j--;
    }
    all[j] = t;
  }
  if (n == 0) {
    snprintf(err, errcap, "the filter is empty");
    free(all);
    return NULL;
  }
  ofilterset *s = calloc(1, sizeof *s);
  s->f = malloc(n * sizeof *s->f);
  for (size_t i = 0; i < n; i++) {
    if (s->n && strncmp(all[i].canonical, s->f[s->n - 1].canonical,
                        strlen(s->f[s->n - 1].canonical)) == 0)
      continue;
    s->f[s->n++] = all[i];
  }
  free(all);
  s->fast = malloc((s->n ? s->n : 1) * sizeof *s->fast);
  for (size_t i = 0; i < s->n; i++) {
    if (is_fast(&s->f[i])) s->fast[s->nfast++] = (uint32_t)i;
    else s->needs_checksum |= needs_checksum(&s->f[i]);
  }
  /* narrowest_acceptable(TARGET_FALSE_HITS = 0.01) over widths 16..24 */
  for (uint32_t bits = 16; bits <= 24; bits++) {
    uint64_t *w = build_index(s, bits);
    if (!w) continue;
    free(s->words);
    s->words = w;
    s->index_bits = bits;
    if (occupancy(w, bits) <= 0.01) break;
  }
  build_buckets(s);
  return s;
}

void ofilter_free(ofilterset *s) {
  if (!s) return;
  for (size_t b = 0; b < s->nbuckets; b++) {
    free(s->buckets[b].keys);
    free(s->buckets[b].owners);
  }
  free(s->buckets);
  free(s->words);
  free(s->fast);
  free(s->f);
  free(s);
}

size_t ofilter_count(const ofilterset *s) { return s->n; }
const char *ofilter_canonical(const ofilterset *s, size_t i) { return s->f[i].canonical; }
int ofilter_needs_checksum(const ofilterset *s) { return s->needs_checksum; }
/* Whether every filter is a literal prefix decidable from the key alone. */
int ofilter_all_fast(const ofilterset *s) { return s->nfast == s->n; }

uint32_t ofilter_index(const ofilterset *s, uint32_t want_bits, uint64_t *words, size_t nwords) {
  if (want_bits == 0) {
    if (!s->index_bits) return 0;
    size_t need = (((size_t)1 << s->index_bits) + 63) / 64;
    if (words && nwords >= need) memcpy(words, s->words, need * 8);
    return s->index_bits;
  }
  uint64_t *w = build_index(s, want_bits);
  if (!w) return 0;
  size_t need = (((size_t)1 << want_bits) + 63) / 64;
  if (words && nwords >= need) memcpy(words, w, need * 8);
  free(w);
  return want_bits;
}

int ofilter_match_address(const ofilterset *s, const uint8_t addr[35]) {
  for (size_t i = 0; i < s->n; i++)
    if (matches_address(&s->f[i], addr)) return (int)i;
  return -1;
}

/* BitmapIndex::may_match_head (src/filter.rs:2321-2604) for a key's leading 64
 * bits, big-endian; 1 when the set has no index (nothing is ruled out). */
int ofilter_may_match_head(const ofilterset *s, uint64_t head) {
  if (!s->index_bits) return 1;
  const uint64_t sl = head >> (64 - s->index_bits);
  return (int)((s->words[sl >> 6] >> (sl & 63)) & 1);
}

int ofilter_match_key(const ofilterset *s, const uint8_t packed[32]) {
  /* FilterSet::match_literals (src/filter.rs:2194-2217) */
  if (s->index_bits) {
    size_t sl = slot_of(s->index_bits, packed);
    if (!((s->words[sl >> 6] >> (sl & 63)) & 1)) return -1;
  }
  if (!s->nbuckets) {
    for (size_t k = 0; k < s->nfast; k++)
      if (matches_key(&s->f[s->fast[k]], packed)) return (int)s->fast[k];
    return -1;
  }
  uint64_t head = head64(packed);
  int best = -1;
  for (size_t b = 0; b < s->nbuckets; b++) {
    const obucket *bk = &s->buckets[b];
    uint64_t probe = head & bk->mask;
    size_t lo = 0, hi = bk->n;
    while (lo < hi) {
      size_t mid = (lo + hi) / 2;
      if (bk->keys[mid] < probe) lo = mid + 1;
      else hi = mid;
    }
    for (size_t i = lo; i < bk->n && bk->keys[i] == probe; i++) {
      if (matches_key(&s->f[bk->owners[i]], packed)) {
        if (best < 0 || (int)bk->owners[i] < best) best = (int)bk->owners[i];
      }
    }
    if (best >= 0) return best;
  }
  return best;
}
