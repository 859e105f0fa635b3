// Host-side filter set (see og_filters.hpp).  Semantics follow src/filter.rs:
//   parsing and validation         Pattern::parse :298-351, parse_symbols :414-494,
//                                  check_reachable_tail :570-601, Filter::parse :649-684
//   dedupe / absorption            FilterSet::new :1326-1338 (sorted by canonical
//                                  text; a filter whose text starts with the previous
//                                  kept one's is dropped, :873-882)
//   prefix bitmap                  BitmapIndex :2321-2604, narrowest_acceptable :2279-2299
//   exact lookup                   LengthBucket :2645-2701, match_literals :2194-2217
//   regex: form                    Pattern::parse_regex :353-375 (the body after
//                                  the first colon, trimmed), regexes held apart
//                                  and never absorbed (:869-876, 1344-1358)
#include "og_filters.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>

namespace og {

namespace {

const char kAlphabet[] = "abcdefghijklmnopqrstuvwxyz234567";

int symbol_value(char c) {
  c = (char)std::tolower((unsigned char)c);
  if (c >= 'a' && c <= 'z') return c - 'a';
  if (c >= '2' && c <= '7') return c - '2' + 26;
  return -1;
}

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// Symbol m of a big-endian bit string: bits [5m, 5m+5) (src/filter.rs:259-267).
inline int symbol_at(const uint8_t *bytes, int nbytes, int position) {
  int bit = position * 5, byte = bit / 8, shift = bit % 8;
  unsigned hi = (unsigned)bytes[byte] << 8;
  unsigned lo = byte + 1 < nbytes ? bytes[byte + 1] : 0;
  return (int)(((hi | lo) >> (11 - shift)) & 31u);
}

inline uint64_t head64(const uint8_t *p) {
  uint64_t w = 0;
  for (int i = 0; i < 8; ++i) w = (w << 8) | p[i];
  return w;
}

inline size_t slot_of(uint32_t bits, const uint8_t *packed) {
  return (size_t)(head64(packed) >> (64 - bits));
}

}  // namespace

std::string parse_filter(const std::string &raw_in, Filter &out) {
  std::string raw = trim(raw_in);
  out = Filter();
  out.text = raw;
  std::string body;
  size_t colon = raw.find(':');
  if (colon != std::string::npos) {
    std::string name;
    for (char c : raw.substr(0, colon))
      if (!std::isspace((unsigned char)c)) name += (char)std::tolower((unsigned char)c);
    if (name == "regex") {
      out.anchor = kAnywhere;
      out.literal = false;
      out.canonical = trim(raw.substr(colon + 1));
      const std::string e = Regex::compile(out.canonical, out.regex);
      if (!e.empty()) return e;
      if (out.regex->anchored()) out.anchor = kStart;
      return "";
    }
    if (name == "prefix") out.anchor = kStart;
    else if (name == "contains") out.anchor = kAnywhere;
    else if (name == "suffix") out.anchor = kEnd;
    else return "unknown form \"" + name + "\": expected prefix, contains, suffix or regex";
    body = trim(raw.substr(colon + 1));
  } else {
    out.anchor = kStart;
    size_t i = 0;
    while (i < raw.size() && raw[i] == '^') ++i;
    body = raw.substr(i);
  }
  for (size_t i = 0; i < body.size();) {
    char c = body[i];
    if (c == '?') {
      out.sets.push_back(0xffffffffu);
      ++i;
    } else if (c == '[') {
      size_t end = body.find(']', i);
      if (end == std::string::npos) return "unbalanced [ or ] at position " + std::to_string(i);
      uint32_t set = 0;
      for (size_t j = i + 1; j < end;) {
        int first = symbol_value(body[j]);
        if (first < 0) return "invalid character at position " + std::to_string(j);
        if (j + 2 < end && body[j + 1] == '-') {
          int last = symbol_value(body[j + 2]);
          if (last < 0) return "invalid character at position " + std::to_string(j + 2);
          if (last < first) return "empty range at position " + std::to_string(j);
          for (int v = first; v <= last; ++v) set |= 1u << v;
          j += 3;
        } else {
          set |= 1u << first;
          ++j;
        }
      }
      if (!set) return "empty character class at position " + std::to_string(i);
      out.sets.push_back(set);
      i = end + 1;
    } else if (c == ']') {
      return "unbalanced [ or ] at position " + std::to_string(i);
    } else {
      int v = symbol_value(c);
      if (v < 0) return std::string("invalid character '") + c + "' at position " + std::to_string(i);
      out.sets.push_back(1u << v);
      ++i;
    }
  }
  const int len = (int)out.sets.size();
  if (len == 0) return "the filter is empty";
  if (len > kAddressLen)
    return "the filter is " + std::to_string(len) + " symbols long, an address has only 56";
  if (std::all_of(out.sets.begin(), out.sets.end(), [](uint32_t s) { return s == 0xffffffffu; }))
    return "the pattern fixes no symbol, so every address would match it";
  if (out.anchor == kEnd) {
    if (!(out.sets[len - 1] & (1u << symbol_value('d'))))
      return "no address can end this way: the last symbol must be 'd'";
    const uint32_t tail = (1u << symbol_value('a')) | (1u << symbol_value('i')) |
                          (1u << symbol_value('q')) | (1u << symbol_value('y'));
    if (len >= 2 && !(out.sets[len - 2] & tail))
      return "no address can end this way: the penultimate symbol must be one of aiqy";
  }
  for (uint32_t s : out.sets) {
    out.canonical += kAlphabet[__builtin_ctz(s)];
    out.literal = out.literal && __builtin_popcount(s) == 1;
  }
  const int head = std::min(len, kKeyOnlySymbols);
  unsigned acc = 0, bits = 0;
  int nb = 0;
  for (int i = 0; i < head; ++i) {
    acc = (acc << 5) | (unsigned)__builtin_ctz(out.sets[i]);
    bits += 5;
    if (bits >= 8) {
      bits -= 8;
      out.bytes[nb++] = (uint8_t)(acc >> bits);
    }
  }
  if (bits) out.bytes[nb++] = (uint8_t)(acc << (8 - bits));
  const int total = head * 5;
  out.nbytes = (total + 7) / 8;
  const int rem = total % 8;
  out.mask = rem == 0 ? 0xff : (uint8_t)(0xff << (8 - rem));
  out.bytes[out.nbytes - 1] &= out.mask;
  return "";
}

bool Filter::matches_key(const uint8_t packed[32]) const {
  if (regex) return false;  // decided on the address
  const int last = nbytes - 1;
  return std::memcmp(packed, bytes, (size_t)last) == 0 && (packed[last] & mask) == bytes[last];
}

void address_text56(const uint8_t addr[35], char out[56]) {
  for (int m = 0; m < kAddressLen; ++m) out[m] = kAlphabet[symbol_at(addr, 35, m)];
}

bool Filter::matches_address(const uint8_t addr[35]) const {
  if (regex) {
    char text[kAddressLen];
    address_text56(addr, text);
    return regex->find(text, kAddressLen);
  }
  const int len = (int)sets.size();
  int first = 0, last = 0;
  if (anchor == kEnd) first = last = kAddressLen - len;
  else if (anchor == kAnywhere) last = kAddressLen - len;
  for (int s = first; s <= last; ++s) {
    bool ok = true;
    for (int i = 0; i < len && ok; ++i) ok = (sets[i] >> symbol_at(addr, 35, s + i)) & 1u;
    if (ok) return true;
  }
  return false;
}

double bitmap_occupancy(const std::vector<uint64_t> &words, uint32_t bits) {
  uint64_t set = 0;
  for (uint64_t w : words) set += (uint64_t)__builtin_popcountll(w);
  return (double)set / (double)(1ull << bits);
}

std::vector<uint64_t> bitmap_for(const FilterSet &s, uint32_t bits) {
  std::vector<uint64_t> words(((size_t)1 << bits) / 64 + (bits < 6 ? 1 : 0), 0);
  for (uint32_t i : s.fast) {
    const Filter &f = s.filters[i];
    uint8_t probe[32] = {0};
    std::memcpy(probe, f.bytes, (size_t)f.nbytes);
    const uint32_t prefix_bits = (uint32_t)f.sets.size() * 5;
    const size_t first = slot_of(bits, probe);
    const size_t run = prefix_bits >= bits ? 1 : (size_t)1 << (bits - prefix_bits);
    for (size_t sl = first; sl < first + run; ++sl) words[sl >> 6] |= 1ull << (sl & 63);
  }
  return words;
}

std::string build_filter_set(const std::vector<std::string> &lines, uint32_t index_bits,
                             FilterSet &out) {
  out = FilterSet();
  std::vector<Filter> parsed, regexes;
  for (const std::string &line : lines) {
    if (trim(line).empty()) continue;
    Filter f;
    std::string e = parse_filter(line, f);
    if (!e.empty()) return trim(line) + ": " + e;
    (f.is_regex() ? regexes : parsed).push_back(std::move(f));
  }
  auto by_canonical = [](const Filter &a, const Filter &b) { return a.canonical < b.canonical; };
  std::stable_sort(parsed.begin(), parsed.end(), by_canonical);
  for (Filter &f : parsed) {
    if (!out.filters.empty() && f.canonical.rfind(out.filters.back().canonical, 0) == 0) continue;
    out.filters.push_back(std::move(f));
  }
  // expressions absorb nothing and are absorbed by nothing: their texts say
  // nothing about the addresses they accept
  std::stable_sort(regexes.begin(), regexes.end(), by_canonical);
  for (Filter &f : regexes) out.filters.push_back(std::move(f));
  for (uint32_t i = 0; i < out.filters.size(); ++i) {
    const Filter &f = out.filters[i];
    (f.is_regex() ? out.regexes : f.is_fast() ? out.fast : out.general).push_back(i);
  }
  if (!out.fast.empty()) {
    if (index_bits) {
      if (index_bits >= 8 && index_bits <= 32) {
        std::vector<uint64_t> w = bitmap_for(out, index_bits);
        if (bitmap_occupancy(w, index_bits) <= 0.25) {
          out.words = std::move(w);
          out.index_bits = index_bits;
        }
      }
    } else {
      for (uint32_t bits = 16; bits <= 24; ++bits) {
        std::vector<uint64_t> w = bitmap_for(out, bits);
        const double occ = bitmap_occupancy(w, bits);
        if (occ > 0.25) continue;
        out.words = std::move(w);
        out.index_bits = bits;
        if (occ <= 0.01) break;
      }
    }
    // One bucket per distinct length (src/filter.rs:1884-1910); built whatever
    // the count, since the device's exact table is made from them.
    for (int sym = 1; sym <= kKeyOnlySymbols; ++sym) {
      std::vector<std::pair<uint64_t, uint32_t>> e;
      const uint32_t bits = (uint32_t)std::min(sym * 5, 64);
      const uint64_t mask = bits == 64 ? ~0ull : ~0ull << (64 - bits);
      for (uint32_t i : out.fast) {
        const Filter &f = out.filters[i];
        if ((int)f.sets.size() != sym) continue;
        uint8_t padded[32] = {0};
        std::memcpy(padded, f.bytes, (size_t)f.nbytes);
        e.push_back({head64(padded) & mask, i});
      }
      if (e.empty()) continue;
      std::sort(e.begin(), e.end());
      Bucket b;
      b.symbols = sym;
      b.mask = mask;
      for (auto &kv : e) {
        b.keys.push_back(kv.first);
        b.owners.push_back(kv.second);
      }
      out.buckets.push_back(std::move(b));
    }
  }
  return "";
}

int FilterSet::match_key(const uint8_t packed[32]) const {
  if (index_bits) {
    const size_t sl = slot_of(index_bits, packed);
    if (!((words[sl >> 6] >> (sl & 63)) & 1)) return -1;
  }
  const uint64_t head = head64(packed);
  int best = -1;
  for (const Bucket &b : buckets) {
    const uint64_t probe = head & b.mask;
    auto it = std::lower_bound(b.keys.begin(), b.keys.end(), probe);
    for (; it != b.keys.end() && *it == probe; ++it) {
      const uint32_t owner = b.owners[(size_t)(it - b.keys.begin())];
      if (filters[owner].matches_key(packed) && (best < 0 || (int)owner < best)) best = (int)owner;
    }
    if (best >= 0) return best;
  }
  return best;
}

int FilterSet::match_address(const uint8_t addr[35]) const {
  for (size_t i = 0; i < filters.size(); ++i)
    if (filters[i].matches_address(addr)) return (int)i;
  return -1;
}

DevPattern dev_pattern(const Filter &f) {
  DevPattern p;
  std::memset(&p, 0, sizeof p);
  p.len = (uint32_t)f.sets.size();
  if (f.anchor == kStart) p.first = p.last = 0;
  else if (f.anchor == kEnd) p.first = p.last = kAddressLen - p.len;
  else {
    p.first = 0;
    p.last = kAddressLen - p.len;
  }
  // Placements whose every symbol sits below 49 are decidable from the key
  // without the sign bit or the checksum (src/filter.rs:120-134).
  const int kl = kKeyOnlySymbols - (int)p.len;
  p.key_last = kl < (int)p.first ? 0xffffffffu : (uint32_t)std::min(kl, (int)p.last);
  for (uint32_t i = 0; i < p.len; ++i) p.sets[i] = f.sets[i];
  return p;
}

}  // namespace og
