// Host-side filter set: the reference's FilterSet (src/filter.rs) for the
// forms the device path serves, compiled into the tables the kernels read.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "og_regex.hpp"

namespace og {

constexpr int kAddressLen = 56;     // src/key.rs:19
constexpr int kKeyOnlySymbols = 49; // src/filter.rs:134

enum Anchor : uint32_t { kStart = 0, kAnywhere = 1, kEnd = 2 };

struct Filter {
  Anchor anchor = kStart;
  std::vector<uint32_t> sets;  // one 32-bit symbol set per position
  std::string canonical;       // lowest symbol of each set (src/filter.rs:336-340)
  std::string text;            // as written
  bool literal = true;
  uint8_t bytes[32] = {0};     // decoded head (<= 49 symbols), padding cleared
  int nbytes = 0;
  uint8_t mask = 0xff;         // significant bits of the last byte
  std::shared_ptr<const Regex> regex;  // the regex: form (no symbol sets)

  bool is_regex() const { return regex != nullptr; }
  bool is_fast() const {
    return !regex && anchor == kStart && literal && (int)sets.size() <= kKeyOnlySymbols;
  }
  bool matches_key(const uint8_t packed[32]) const;
  bool matches_address(const uint8_t addr[35]) const;
};

// One length bucket of the exact table (src/filter.rs:2645-2701).
struct Bucket {
  int symbols;
  uint64_t mask;
  std::vector<uint64_t> keys;   // masked big-endian 64-bit heads, sorted
  std::vector<uint32_t> owners; // index into FilterSet::filters
};

// Device form of a general pattern: the placements it may take and one symbol
// set per position, checked against the 35 address bytes.
struct DevPattern {
  uint32_t len;
  uint32_t first;  // first start position
  uint32_t last;   // last start position
  uint32_t key_last;  // placements first..key_last fit in the key-only span
  uint32_t sets[kAddressLen];
};

struct FilterSet {
  std::vector<Filter> filters;  // sorted, deduplicated, absorbed
  std::vector<uint32_t> fast;   // literal prefixes (indices into filters)
  std::vector<uint32_t> general;  // suffix, contains, class and long prefixes
  std::vector<uint32_t> regexes;  // regex: filters, after all the others
  uint32_t index_bits = 0;      // 0: no index (src/filter.rs:2321)
  std::vector<uint64_t> words;  // the bitmap
  std::vector<Bucket> buckets;

  int match_key(const uint8_t packed[32]) const;
  int match_address(const uint8_t addr[35]) const;
};

// Returns an empty string on success, else the message.
std::string parse_filter(const std::string &raw, Filter &out);
std::string build_filter_set(const std::vector<std::string> &lines, uint32_t index_bits, FilterSet &out);
// A bitmap over the literal prefixes at `bits` without the occupancy cap: the
// device's prefilter when the reference would build none.
std::vector<uint64_t> bitmap_for(const FilterSet &s, uint32_t bits);
double bitmap_occupancy(const std::vector<uint64_t> &words, uint32_t bits);
DevPattern dev_pattern(const Filter &f);
// The 56 symbols of an address as text.
void address_text56(const uint8_t addr[35], char out[56]);

}  // namespace og
