// Host side of the library: curve constants, the offset table, key derivation
// (src/key.rs) and the filter-set entry points of the C ABI.
#include "og_host.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "fe10.hpp"
#include "keccak.hpp"
#include "og_devfilter.hpp"
#include "og_scan.hpp"
#include "sha512.hpp"

struct og_filterset {
  og::FilterSet set;
  // the device tables, compiled once on the first host-side use
  std::once_flag once;
  og::DevFilterHost tables;
  std::string tables_err;

  const og::DevFilterHost &host_tables() {
    std::call_once(once, [this] { tables_err = tables.compile(set); });
    return tables;
  }
};

namespace og {

namespace {

Fe fe_small(uint32_t v) {
  Fe r = fe_zero();
  r.v[0] = v;
  return r;
}

void put_err(char *err, size_t cap, const std::string &msg) {
  if (err && cap) {
    size_t n = msg.size() < cap - 1 ? msg.size() : cap - 1;
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

// The base point's y is 4/5; x is the even root of (y^2 - 1)/(d y^2 + 1).
// Both are the published RFC 8032 constants, computed rather than typed.
Fe sqrt_ratio_even(const Fe &u, const Fe &v) {
  // x = u v^3 (u v^7)^((p-5)/8), then fix by sqrt(-1) when needed (RFC 8032 5.1.3)
  Fe v3 = fe_mul(fe_sq(v), v);
  Fe v7 = fe_mul(fe_sq(v3), v);
  Fe t = fe_mul(u, v7);
  // (p-5)/8 = 2^252 - 3: t^(2^252-3) via the chain of t^(2^250-1)
  Fe z2 = fe_sq(t);
  Fe z8 = fe_sqn(z2, 2);
  Fe z9 = fe_mul(t, z8);
  Fe z11 = fe_mul(z2, z9);
  Fe z22 = fe_sq(z11);
  Fe z5 = fe_mul(z9, z22);
  Fe z10 = fe_mul(fe_sqn(z5, 5), z5);
  Fe z20 = fe_mul(fe_sqn(z10, 10), z10);
  Fe z40 = fe_mul(fe_sqn(z20, 20), z20);
  Fe z50 = fe_mul(fe_sqn(z40, 10), z10);
  Fe z100 = fe_mul(fe_sqn(z50, 50), z50);
  Fe z200 = fe_mul(fe_sqn(z100, 100), z100);
  Fe z250 = fe_mul(fe_sqn(z200, 50), z50);
  Fe pw = fe_mul(fe_sqn(z250, 2), t);  // t^(2^252 - 3)
  Fe x = fe_mul(fe_mul(u, v3), pw);
  Fe vx2 = fe_mul(v, fe_sq(x));
  if (!fe_is_zero(fe_sub(vx2, u))) {
    // sqrt(-1) = 2^((p-1)/4)
    Fe two = fe_small(2);
    Fe i = two;
    // 2^((p-1)/4) = 2^(2^253 - 5): square-and-multiply over the exponent bits
    // 2^253 - 5 = 0b1...1011 (253 bits)
    Fe acc = fe_one();
    for (int bit = 252; bit >= 0; --bit) {
      acc = fe_sq(acc);
      const bool set = !(bit == 2);  // 2^253 - 5 = all ones except bit 2
      if (set) acc = fe_mul(acc, two);
    }
    i = acc;
    x = fe_mul(x, i);
  }
  uint8_t b[32];
  fe_to_bytes(b, x);
  if (b[0] & 1) x = fe_sub(fe_zero(), x);
  return fe_canon(x);
}

}  // namespace

HostCurve::HostCurve() {
  // d = -121665 / 121666 (RFC 8032 5.1)
  d = fe_canon(fe_mul(fe_sub(fe_zero(), fe_small(121665)), fe_invert(fe_small(121666))));
  two_d = fe_canon(fe_add(d, d));
  Fe y = fe_canon(fe_mul(fe_small(4), fe_invert(fe_small(5))));
  Fe y2 = fe_sq(y);
  Fe x = sqrt_ratio_even(fe_sub(y2, fe_one()), fe_add(fe_mul(d, y2), fe_one()));
  base.X = x;
  base.Y = y;
  base.Z = fe_one();
  base.T = fe_canon(fe_mul(x, y));
}

Cached HostCurve::eight_cached() const {
  Point p = pt_dbl(pt_dbl(pt_dbl(base)));
  return pt_to_cached(p, two_d);
}

std::vector<uint32_t> HostCurve::offset_table(uint32_t half) const {
  std::vector<uint32_t> out((size_t)half * 24);
  if (!half) return out;
  const Cached e8 = eight_cached();
  std::vector<Point> pts(half);
  Point q = pt_add(pt_identity(), e8);
  for (uint32_t m = 0; m < half; ++m) {
    pts[m] = q;
    q = pt_add(q, e8);
  }
  // one inversion for the whole table (Montgomery's trick)
  std::vector<Fe> pre(half);
  Fe acc = fe_one();
  for (uint32_t m = 0; m < half; ++m) {
    pre[m] = acc;
    acc = fe_mul(acc, pts[m].Z);
  }
  Fe inv = fe_invert(acc);
  for (uint32_t m = half; m-- > 0;) {
    Fe zinv = fe_mul(inv, pre[m]);
    inv = fe_mul(inv, pts[m].Z);
    Fe x = fe_canon(fe_mul(pts[m].X, zinv));
    Fe y = fe_canon(fe_mul(pts[m].Y, zinv));
    Fe dxy = fe_canon(fe_mul(d, fe_mul(x, y)));
    for (int l = 0; l < 8; ++l) {
      out[(size_t)m * 24 + l] = x.v[l];
      out[(size_t)m * 24 + 8 + l] = y.v[l];
      out[(size_t)m * 24 + 16 + l] = dxy.v[l];
    }
  }
  return out;
}

std::vector<uint32_t> HostCurve::offset_table10(uint32_t half) const {
  const std::vector<uint32_t> t8 = offset_table(half);
  std::vector<uint32_t> out((size_t)half * kTableWords);
  for (uint32_t m = 0; m < half; ++m) {
    Fe c[kTableTerms];
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 8; ++l) c[i].v[l] = t8[(size_t)m * 24 + 8 * i + l];
    offset_terms(c);
    for (uint32_t i = 0; i < kTableTerms; ++i)
      put_fe10(&out[(size_t)m * kTableWords + 10 * i], c[i]);
  }
  return out;
}

SetupArgs HostCurve::setup_args(const uint8_t start_xyzt[128], uint32_t *state, uint32_t threads,
                                uint32_t first, uint32_t count, uint32_t half, uint64_t k0) const {
  SetupArgs sa;
  std::memset(&sa, 0, sizeof sa);
  sa.state = state;
  sa.threads = threads;
  sa.first = first;
  sa.count = count;
  sa.half = half;
  sa.k0 = k0;
  for (int c = 0; c < 4; ++c) {
    Fe v = fe_from_bytes(start_xyzt + 32 * c);
    for (int l = 0; l < 8; ++l) sa.start[8 * c + l] = v.v[l];
  }
  const Cached eight = eight_cached();
  for (int l = 0; l < 8; ++l) {
    sa.eight[l] = eight.ypx.v[l];
    sa.eight[8 + l] = eight.ymx.v[l];
    sa.eight[16 + l] = eight.Z.v[l];
    sa.eight[24 + l] = eight.t2d.v[l];
    sa.two_d[l] = two_d.v[l];
  }
  return sa;
}

void put_fe10(uint32_t out[10], const Fe &x) {
  const Fe10 r = fe10_from(x);
  for (int l = 0; l < 10; ++l) out[l] = r.v[l];
}

Cached1 HostCurve::multiple_of_eight_cached1(uint64_t m) const {
  Point p = pt_mul_u64(m, eight_cached());
  Fe zinv = fe_invert(p.Z);
  Fe x = fe_mul(p.X, zinv), y = fe_mul(p.Y, zinv);
  Cached1 c;
  c.ypx = fe_canon(fe_add(y, x));
  c.ymx = fe_canon(fe_sub(y, x));
  c.t2d = fe_canon(fe_mul(two_d, fe_mul(x, y)));
  return c;
}

Point HostCurve::scalar_base_mult(const uint8_t scalar[32]) const {
  const Cached bc = pt_to_cached(base, two_d);
  Point acc = pt_identity();
  for (int i = 255; i >= 0; --i) {
    acc = pt_dbl(acc);
    if ((scalar[i / 8] >> (i % 8)) & 1) acc = pt_add(acc, bc);
  }
  return acc;
}

void HostCurve::pack(const Point &p, uint8_t out[32]) const {
  Fe zinv = fe_invert(p.Z);
  Fe y = fe_mul(p.Y, zinv), x = fe_mul(p.X, zinv);
  fe_to_bytes(out, y);
  uint8_t xb[32];
  fe_to_bytes(xb, x);
  out[31] ^= (uint8_t)((xb[0] & 1) << 7);
}

const FilterSet *filterset_of(const og_filterset *s) { return s ? &s->set : nullptr; }

static const HostCurve &curve() {
  static const HostCurve c;
  return c;
}

}  // namespace og

using namespace og;

extern "C" {

og_filterset *og_filters_parse(const char *text, uint32_t index_bits, char *err, size_t errcap) {
  if (!text) {
    put_err(err, errcap, "null filter text");
    return nullptr;
  }
  std::vector<std::string> lines;
  std::string cur;
  for (const char *p = text; *p; ++p) {
    if (*p == '\n') {
      lines.push_back(cur);
      cur.clear();
    } else {
      cur += *p;
    }
  }
  lines.push_back(cur);
  og_filterset *s = new og_filterset();
  std::string e = build_filter_set(lines, index_bits, s->set);
  if (!e.empty() || s->set.filters.empty()) {
    put_err(err, errcap, e.empty() ? "no filters" : e);
    delete s;
    return nullptr;
  }
  return s;
}

void og_filters_free(og_filterset *s) { delete s; }

size_t og_filters_count(const og_filterset *s) { return s ? s->set.filters.size() : 0; }

uint32_t og_filters_index_words(const og_filterset *s, uint64_t *words, size_t cap) {
  if (!s || !s->set.index_bits) return 0;
  if (words && cap >= s->set.words.size())
    std::memcpy(words, s->set.words.data(), s->set.words.size() * 8);
  return s->set.index_bits;
}

int og_filters_match_key(const og_filterset *s, const uint8_t packed[32]) {
  return s ? s->set.match_key(packed) : -1;
}

int og_filters_match_address(const og_filterset *s, const uint8_t address[35]) {
  return s ? s->set.match_address(address) : -1;
}

int og_filters_scan_pair_host(const og_filterset *s, const uint8_t a[32], const uint8_t b[32],
                              int key_hit[2], int need[2]) {
  if (!s || !a || !b || !key_hit || !need) return -1;
  const DevFilterHost &h = const_cast<og_filterset *>(s)->host_tables();
  if (!s->tables_err.empty() || !(h.mode & kModePattern)) return -1;
  const uint8_t *keys[2] = {a, b};
  uint32_t kw[2][8];
  bool kh[2], nd[2], maybe[2];
  for (int i = 0; i < 2; ++i) {
    Fe c;
    for (int l = 0; l < 8; ++l)
      c.v[l] = (uint32_t)keys[i][4 * l] | ((uint32_t)keys[i][4 * l + 1] << 8) |
               ((uint32_t)keys[i][4 * l + 2] << 16) | ((uint32_t)keys[i][4 * l + 3] << 24);
    key_words(c, kw[i]);
    scan_key(kw[i], h.banks.data(), (uint32_t)h.banks.size(), kh[i], nd[i], maybe[i]);
  }
  const uint32_t nre = (uint32_t)(h.regex.size() / 3);
  if (nre) {
    for (int i = 0; i < 2; ++i) maybe[i] = (maybe[i] || h.regex_every) && !kh[i];
    if (maybe[0] || maybe[1]) {
      const uint32_t K = h.first_states <= kLdsStates ? h.first_states : 0u;
      const uint32_t r =
          h.hash_key ? regex_key2_prefiltered(kw[0], kw[1], h.regex.data(), h.dfa.data(),
                                              h.live.data(), nre, h.hash.data(), h.hash_lmax,
                                              h.hash_ws)
                      : regex_key2(kw[0], kw[1], h.regex.data(), h.dfa.data(), h.live.data(),
                                   nre, h.dfa.data() + h.regex[0], K);
      const uint32_t ri[2] = {maybe[0] ? (r & 3u) : 0u, maybe[1] ? (r >> 2) : 0u};
      for (int i = 0; i < 2; ++i) {
        kh[i] = kh[i] || ri[i] == 2;
        nd[i] = nd[i] || ri[i] == 1;
      }
    }
  }
  for (int i = 0; i < 2; ++i) {
    key_hit[i] = kh[i];
    need[i] = nd[i] && !kh[i];
  }
  return 0;
}

int og_filters_address_window_host(const og_filterset *s, const uint8_t address[35], int *window,
                                   int *match) {
  if (!s || !address || !window || !match) return -1;
  const DevFilterHost &h = const_cast<og_filterset *>(s)->host_tables();
  if (!s->tables_err.empty() || h.regex.empty()) return -1;
  uint32_t w[9];  // big-endian words, as the address kernels form them
  for (int i = 0; i < 8; ++i)
    w[i] = ((uint32_t)address[4 * i] << 24) | ((uint32_t)address[4 * i + 1] << 16) |
           ((uint32_t)address[4 * i + 2] << 8) | address[4 * i + 3];
  w[8] = ((uint32_t)address[32] << 24) | ((uint32_t)address[33] << 16) |
         ((uint32_t)address[34] << 8);
  *window = !h.hash_addr || window_hit<56, 9>(w, h.hash.data(), h.hash_ws);
  *match = regex_address(w, h.regex.data(), h.dfa.data(), h.dfa_end.data(),
                         (uint32_t)(h.regex.size() / 3));
  return 0;
}

size_t og_filters_describe(const og_filterset *s, char *out, size_t cap) {
  if (!s) return 0;
  const FilterSet &fs = s->set;
  std::string t;
  if (!fs.fast.empty())
    t += "prefix table: " + std::to_string(fs.fast.size()) + " literal prefix(es)" +
         (fs.index_bits ? ", bitmap of 2^" + std::to_string(fs.index_bits) + " bits" : "") + "\n";
  const DevFilterHost &h = const_cast<og_filterset *>(s)->host_tables();
  const std::string &e = s->tables_err;
  if (!h.dictionaries.empty()) {
    size_t states = 0;
    for (const auto &d : h.dictionaries) states += d->dfa_states();
    t += (h.dictionaries.size() == 1 ? "dictionary automaton: "
                                     : "dictionary automata: " +
                                           std::to_string(h.dictionaries.size()) + " of ") +
         std::to_string(states) + " states over " + std::to_string(h.dictionary_forms) +
         " general form(s), run on every candidate; the checksum only where a match can still "
         "end in it\n";
    if (!h.hash.empty()) {
      char pct[16];
      std::snprintf(pct, sizeof pct, "%.2f", h.window_share * 100.0);
      const std::string form = "two 2^18-bit maps over " + std::to_string(h.hash_ws / 8) +
                               "-symbol windows at every " +
                               (h.hash_ws % 8 == 1 ? std::string("")
                                                   : std::to_string(h.hash_ws % 8) + "th ") +
                               "placement, in LDS; ";
      if (h.hash_key)
        t += "window prefilter: " + form + "the automaton walks symbols " +
             std::to_string(50 - h.hash_lmax) +
             "..48 for every candidate, and a key whose window hits (about " + pct +
             "%) goes to the address stage\n";
      else
        t += "window prefilter in the address stage: " + form +
             "the whole walk runs only in waves where a window of the 56 symbols hits "
             "(about " + pct + "% of waves)\n";
    }
  }
  if (h.park_all) {
    char pct[16];
    std::snprintf(pct, sizeof pct, "%.1f", h.need * 100.0);
    t += std::string("key stage skipped: ") + pct +
         "% of sample keys can still match in the checksum, so every candidate goes to the "
         "address stage\n";
  }
  const bool dict = !h.dictionaries.empty();
  if ((!fs.general.empty() && !dict) || (!fs.regexes.empty() && !h.key_proxy)) {
    t += "automaton banks: " + std::to_string(h.banks.size()) + " for " +
         std::to_string(dict ? 0 : fs.general.size()) + " general form(s)";
    if (!fs.regexes.empty()) t += " and the proxies of " + std::to_string(fs.regexes.size()) + " expression(s)";
    t += e.empty() ? "\n" : " (" + e + ")\n";
    size_t lits = 0;
    for (const ScanBank &b : h.banks) lits += b.lit ? 1 : 0;
    if (lits)
      t += "literal compare: " + std::to_string(lits) +
           " bank(s) of at most two literals, compared at their placements in the key pass\n";
  }
  if (h.key_proxy)
    t += "key tables: " + std::to_string(h.proxy_set.fast.size()) +
         " literal prefix(es) of the expressions, each hit confirmed by their automata\n";
  for (uint32_t i : fs.regexes) {
    const Regex &re = *fs.filters[i].regex;
    t += re.describe();
    t += re.key_decidable() ? "; decided within the first 49 symbols"
                            : "; an open candidate costs a SHA3-256";
    t += "\n";
  }
  if (h.proxyless)
    t += "an expression has no proxy in the banks: its automaton runs on every candidate\n";
  if (out && cap) {
    const size_t n = std::min(cap - 1, t.size());
    std::memcpy(out, t.data(), n);
    out[n] = 0;
  }
  return t.size();
}

int og_filters_exact_table_host(const og_filterset *s, const uint8_t packed[32], int *hit) {
  if (!s || !packed || !hit) return -1;
  const DevFilterHost &h = const_cast<og_filterset *>(s)->host_tables();
  if (!s->tables_err.empty() || !(h.mode & kModeBitmap) || h.buckets.empty()) return -1;
  const uint32_t hi = ((uint32_t)packed[0] << 24) | ((uint32_t)packed[1] << 16) |
                      ((uint32_t)packed[2] << 8) | packed[3];
  const uint32_t lo = ((uint32_t)packed[4] << 24) | ((uint32_t)packed[5] << 16) |
                      ((uint32_t)packed[6] << 8) | packed[7];
  *hit = exact_table(hi, lo, h.buckets.data(), h.keys.data(), h.dirs.data(),
                     (uint32_t)(h.buckets.size() / kBucketWords), h.dir_steps)
             ? 1
             : 0;
  return 0;
}

int og_filters_scan_host(const og_filterset *s, const uint8_t packed[32], int sign,
                         int *key_hit, int *need, int *hit) {
  if (!s || !packed || !key_hit || !need || !hit) return -1;
  // the tables the device reads, compiled the same way
  const DevFilterHost &h = const_cast<og_filterset *>(s)->host_tables();
  if (!s->tables_err.empty() || !(h.mode & kModePattern || h.key_proxy)) return -1;
  const ScanBank *banks = h.banks.data();
  const uint32_t nb = (uint32_t)h.banks.size(), nre = (uint32_t)(h.regex.size() / 3);
  Fe c;
  for (int l = 0; l < 8; ++l) {
    c.v[l] = (uint32_t)packed[4 * l] | ((uint32_t)packed[4 * l + 1] << 8) |
             ((uint32_t)packed[4 * l + 2] << 16) | ((uint32_t)packed[4 * l + 3] << 24);
  }
  uint32_t kw[8];
  key_words(c, kw);
  // the key stage as the search kernels run it (without the key forms)
  bool kh = false, nd = false, maybe = false;
  if (h.key_proxy) {
    // the expressions' literal prefixes in the key tables, then the automata
    kh = h.proxy_set.match_key(packed) >= 0 &&
         regex_key(kw, h.regex.data(), h.dfa.data(), h.live.data(), nre, h.dfa.data(), 0u) == 2;
  } else {
    scan_key(kw, banks, nb, kh, nd, maybe);
  }
  if (nre && !h.key_proxy) {
    maybe = (maybe || h.regex_every) && !kh;
    if (maybe) {
      // the first automaton in LDS as the kernels hold it, when it fits
      const uint32_t K = h.first_states <= kLdsStates ? h.first_states : 0u;
      const uint32_t r = regex_key(kw, h.regex.data(), h.dfa.data(), h.live.data(), nre,
                                   h.dfa.data() + h.regex[0], K);
      kh |= r == 2;
      nd |= r == 1;
    }
  }
  nd &= !kh;
  // the address stage, on every key
  uint32_t pk[8];
  for (int l = 0; l < 8; ++l) pk[l] = c.v[l];
  pk[7] |= (uint32_t)(sign & 1) << 31;
  uint32_t w[9];
  address_words(kw, (uint32_t)(sign & 1), onion_checksum(pk), w);
  uint32_t r = scan_address(w, banks, nb);
  if (r == 2) {
    // a truncated pattern's kept part matched: the symbol-wise check
    uint8_t addr[35];
    for (int i = 0; i < 35; ++i) addr[i] = (uint8_t)(w[i / 4] >> (24 - 8 * (i % 4)));
    r = 0;
    for (uint32_t g : s->set.general)
      if (s->set.filters[g].matches_address(addr)) r = 1;
  }
  if (r != 1 && nre && regex_address(w, h.regex.data(), h.dfa.data(), h.dfa_end.data(), nre)) r = 1;
  *key_hit = kh;
  *need = nd;
  *hit = r == 1;
  return 0;
}

void og_key_secret_scalar(const uint8_t seed[32], uint8_t sk[32]) {
  uint8_t d[64];
  sha512(seed, 32, d);
  std::memcpy(sk, d, 32);
  sk[0] &= 248;
  sk[31] &= 63;
  sk[31] |= 64;
}

void og_key_expanded_secret(const uint8_t seed[32], uint8_t out[64]) {
  sha512(seed, 32, out);
  out[0] &= 248;
  out[31] &= 63;
  out[31] |= 64;
}

void og_key_start_point(const uint8_t seed[32], uint8_t xyzt[128]) {
  uint8_t sk[32];
  og_key_secret_scalar(seed, sk);
  Point p = curve().scalar_base_mult(sk);
  fe_to_bytes(xyzt, p.X);
  fe_to_bytes(xyzt + 32, p.Y);
  fe_to_bytes(xyzt + 64, p.Z);
  fe_to_bytes(xyzt + 96, p.T);
}

void og_key_public_from_scalar(const uint8_t scalar[32], uint8_t pk[32]) {
  curve().pack(curve().scalar_base_mult(scalar), pk);
}

int og_key_expanded_secret_at_offset(const uint8_t seed[32], uint64_t offset, uint8_t out[64]) {
  og_key_expanded_secret(seed, out);
  unsigned __int128 carry = offset;
  for (int i = 0; i < 32 && carry; ++i) {
    carry += out[i];
    out[i] = (uint8_t)carry;
    carry >>= 8;
  }
  return (out[0] & 248) == out[0] && (((out[31] & 63) | 64) == out[31]);
}

void og_key_address_bytes(const uint8_t pk[32], uint8_t out[35]) {
  uint32_t w[8];
  for (int l = 0; l < 8; ++l)
    w[l] = (uint32_t)pk[4 * l] | ((uint32_t)pk[4 * l + 1] << 8) | ((uint32_t)pk[4 * l + 2] << 16) |
           ((uint32_t)pk[4 * l + 3] << 24);
  const uint32_t chk = onion_checksum(w);
  std::memcpy(out, pk, 32);
  out[32] = (uint8_t)chk;
  out[33] = (uint8_t)(chk >> 8);
  out[34] = 0x03;
}

void og_key_address(const uint8_t pk[32], char out[57]) {
  static const char alpha[] = "abcdefghijklmnopqrstuvwxyz234567";
  uint8_t b[35];
  og_key_address_bytes(pk, b);
  unsigned acc = 0, bits = 0;
  int o = 0;
  for (int i = 0; i < 35; ++i) {
    acc = (acc << 8) | b[i];
    bits += 8;
    while (bits >= 5) {
      bits -= 5;
      out[o++] = alpha[(acc >> bits) & 31];
    }
  }
  out[o] = 0;
}

void og_run_block_seed(const uint8_t root[32], uint64_t index, uint8_t out[32]) {
  uint8_t buf[58], d[64];
  std::memcpy(buf, "onion-gen block v1", 18);
  std::memcpy(buf + 18, root, 32);
  for (int i = 0; i < 8; ++i) buf[50 + i] = (uint8_t)(index >> (8 * i));
  sha512(buf, sizeof buf, d);
  std::memcpy(out, d, 32);
}

}  // extern "C"
