// Host side: a FilterSet compiled into the tables a search kernel reads, and
// their device copies.  Shared by the single-search engine and the slot
// engine.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "og_filters.hpp"
#include "og_kernels.hpp"

namespace og {

// The general filters of a set, then its expressions' proxies, packed into
// automaton banks (og_scan.hpp): first fit, each pattern on consecutive bits
// of one bank.  An expression whose proxies do not fit even shortened gets
// none, and `every` says so: its automaton then runs on every candidate.
// Empty string on success.
inline std::string build_banks(const FilterSet &fs, std::vector<ScanBank> &out,
                               bool *every = nullptr, bool general = true) {
  out.clear();
  std::vector<int> used;
  // per bank: its literals that may start inside the key (value, mask and
  // info as ScanBank takes them), and whether some other pattern may start
  // there too (then the bank walks every symbol)
  struct Lit {
    uint32_t val, mask, info;
  };
  std::vector<std::vector<Lit>> lits;
  std::vector<bool> walks;
  // 0: placed, 1: no room
  auto place = [&](const std::vector<uint32_t> &sets, Anchor anchor, bool proxy) {
    const int len = (int)sets.size();
    const int kept = std::min(len, kBankBits);
    size_t k = 0;
    while (k < out.size() && used[k] + kept > kBankBits) ++k;
    if (k == out.size()) {
      if (out.size() == (size_t)kMaxBanks) return 1;
      ScanBank b;
      std::memset(&b, 0, sizeof b);
      b.jstart = kAddressLen;
      out.push_back(b);
      used.push_back(0);
      lits.emplace_back();
      walks.push_back(false);
    }
    ScanBank &b = out[k];
    const int off = used[k];
    used[k] += kept;
    if (anchor == kEnd && len <= 7) {
      // starts past symbol 48: only the tail sees it
    } else {
      // a literal of its first (up to 6) symbols; a longer proxy's first six
      // only ask for the automata more often, never less
      bool literal = len >= 1 && len <= (proxy ? 7 : 6);
      const int n = std::min(len, 6);
      uint32_t v = 0;
      for (int i = 0; i < n && literal; ++i) {
        literal = __builtin_popcount(sets[i]) == 1;
        v |= (uint32_t)__builtin_ctz(sets[i] | (1u << 31)) << (27 - 5 * i);
      }
      for (int i = n; i < len && literal; ++i) literal = __builtin_popcount(sets[i]) == 1;
      if (literal) {
        Lit l{v, ~0u << (32 - 5 * n), 0u};
        for (int j = 44; j < 49; ++j)
          if (j + len <= 49) l.info |= 1u << (j - 44);
        if (anchor == kStart) l.info |= kLitStart;
        if (proxy) l.info |= kLitProxy;
        lits[k].push_back(l);
      } else {
        walks[k] = true;
      }
    }
    for (int i = 0; i < kept; ++i)
      for (int c = 0; c < 32; ++c)
        if ((sets[i] >> c) & 1u) b.b[c] |= 1u << (off + i);
    const uint32_t first = 1u << off, last = 1u << (off + kept - 1);
    if (anchor == kStart) b.istart |= first;
    else b.iany |= first;
    if (proxy) b.fmaybe |= last;
    else if (len > kBankBits) b.flong |= last;
    else if (anchor == kEnd) b.fend |= last;
    else b.fany |= last;
    const uint32_t js = anchor == kEnd ? (uint32_t)(kAddressLen - len) : 0u;
    b.jstart = std::min(b.jstart, js);
    return 0;
  };
  for (uint32_t idx : general ? fs.general : std::vector<uint32_t>()) {
    const Filter &f = fs.filters[idx];
    if (place(f.sets, f.anchor, false))
      return "the general filters need more than 8 automaton banks";
  }
  if (every) *every = false;
  for (uint32_t idx : fs.regexes) {
    const Regex &re = *fs.filters[idx].regex;
    int free_bits = (kMaxBanks - (int)out.size()) * kBankBits;
    for (int u : used) free_bits += kBankBits - u;
    bool placed = false;
    // fragmentation can defeat a budget that fits in total: halve and retry
    for (int budget = free_bits; budget > 0 && !placed; budget /= 2) {
      const std::vector<RegexProxy> ps = re.proxies(budget);
      if (ps.empty()) break;
      const std::vector<ScanBank> out0 = out;
      const std::vector<int> used0 = used;
      placed = true;
      for (const RegexProxy &p : ps)
        if (place(p.sets, p.anchored ? kStart : kAnywhere, true)) placed = false;
      if (!placed) {
        out = out0;
        used = used0;
      }
    }
    if (!placed && every) *every = true;
  }
  for (size_t k = 0; k < out.size(); ++k) {
    ScanBank &b = out[k];
    for (int c = 0; c < 32; ++c) {
      b.u51[c >> 4] |= b.b[c];
      b.uall |= b.b[c];
    }
    b.u54 = b.b[0] | b.b[8] | b.b[16] | b.b[24];  // a, i, q, y
    if (!walks[k] && !lits[k].empty() && lits[k].size() <= 2) {
      // literals compared at each placement, about 3 instructions a
      // placement each (one for a start-anchored one) where the automaton
      // takes about 6 a symbol
      b.lit = (uint32_t)lits[k].size();
      for (size_t i = 0; i < lits[k].size(); ++i) {
        b.lit_val[i] = lits[k][i].val;
        b.lit_mask[i] = lits[k][i].mask;
        b.lit_info[i] = lits[k][i].info;
      }
    }
  }
  return "";
}

// The masked compare of one literal prefix (kModePrefix): the decoded head
// bytes and their significant bits as little-endian key words, or'ed in.
inline void prefix_words(const Filter &f, uint32_t want[8], uint32_t mask[8]) {
  for (int i = 0; i < f.nbytes; ++i) {
    const uint8_t m = i == f.nbytes - 1 ? f.mask : 0xff;
    want[i / 4] |= (uint32_t)f.bytes[i] << (8 * (i % 4));
    mask[i / 4] |= (uint32_t)m << (8 * (i % 4));
  }
}

struct DevFilterHost {
  uint32_t mode = 0;
  uint32_t want[8] = {0}, mask[8] = {0};
  uint32_t bits = 0;
  std::vector<uint32_t> bitmap;   // 32-bit halves of the 64-bit words
  std::vector<uint32_t> hash;     // the window prefilter's two maps, or empty
  uint32_t hash_lmax = 0;         // its longest word
  uint32_t hash_ws = 0;           // its form: 8 x window length + stride
  bool hash_key = false;          // the key stage uses it
  bool hash_addr = false;         // the address stage uses it
  std::vector<uint32_t> coarse;   // the bitmap or-folded to coarse_bits bits
  uint32_t coarse_bits = 0;
  std::vector<uint32_t> buckets;  // kBucketWords per distinct prefix length
  std::vector<uint64_t> keys;
  std::vector<uint32_t> dirs;     // the buckets' directories
  uint32_t dir_steps = 0;
  std::vector<uint32_t> pats;     // kPatWords per pattern
  std::vector<ScanBank> banks;    // the automaton over the same patterns
  std::vector<uint32_t> regex;    // per expression: entry offset, first state, live offset
  std::vector<uint16_t> dfa;      // the expressions' automata: next states only
  std::vector<uint32_t> dfa_end;  // per state, bit c: the address may end on
                                  //   symbol c from that state (word off/32 + s
                                  //   of an automaton at entry offset off)
  std::vector<uint32_t> live;     // their tail-liveness bits
  bool regex_every = false;
  bool proxyless = false;         // an expression found no room for its proxies
  // The fraction of sample keys whose key stage leaves a checksum placement
  // possible (need_rate over kNeedSamples), and whether it is high enough
  // that the key stage is skipped: every candidate the key forms do not
  // settle goes to the address stage, whose exact pass over the whole address
  // decides the general forms and expressions alone.
  double need = 0.0;
  bool park_all = false;
  double window_park = 0.0;       // the window prefilter's expected share of
                                  //   keys sent to the address stage
  double window_share = 0.0;      // the same, or (hash_addr) of waves that walk
  // General filters beyond the banks, compiled into automata (a dictionary
  // of substrings and suffixes): one, or several when one would exceed the
  // state limit; run on every candidate.
  std::vector<std::shared_ptr<const Regex>> dictionaries;
  size_t dictionary_forms = 0;
  // A set of key-decidable expressions only: their literal prefixes, which
  // the key tables hold in place of the set's own (key_proxy).
  bool key_proxy = false;
  FilterSet proxy_set;
  // device copies
  uint32_t *d_bitmap = nullptr, *d_buckets = nullptr, *d_pats = nullptr, *d_banks = nullptr;
  uint32_t *d_coarse = nullptr, *d_hash = nullptr;
  uint32_t *d_regex = nullptr, *d_live = nullptr;
  uint16_t *d_dfa = nullptr;
  uint32_t *d_dfa_end = nullptr;
  uint64_t *d_keys = nullptr;
  uint32_t *d_dirs = nullptr;

  size_t device_bytes() const {
    return bitmap.size() * 4 + coarse.size() * 4 + hash.size() * 4 + buckets.size() * 4 + keys.size() * 8 +
           dirs.size() * 4 + pats.size() * 4 +
           banks.size() * sizeof(ScanBank) + regex.size() * 4 + dfa.size() * 2 + dfa_end.size() * 4 +
           live.size() * 4;
  }

  uint32_t first_states = 0;  // states of the first automaton

  void add_automaton(const Regex &re) {
    if (regex.empty()) first_states = re.dfa_states();
    regex.push_back((uint32_t)dfa.size());
    regex.push_back(re.dfa_start());
    regex.push_back((uint32_t)live.size());
    // the walks read next states alone; the end flags go to their own bits
    for (size_t i = 0; i < re.dfa().size(); i += 32) {
      uint32_t ends = 0;
      for (int c = 0; c < 32; ++c) {
        const uint16_t e = re.dfa()[i + c];
        dfa.push_back((uint16_t)(e & ~kRegexEndBit));
        ends |= (e & kRegexEndBit) ? 1u << c : 0u;
      }
      dfa_end.push_back(ends);
    }
    live.insert(live.end(), re.tail_live().begin(), re.tail_live().end());
  }

  // The expressions' automata, in the order of fs.regexes.
  void add_automata(const FilterSet &fs) {
    for (uint32_t i : fs.regexes) add_automaton(*fs.filters[i].regex);
  }

  // The general filters as one expression: an alternation of their symbol
  // sets, anchored as their forms are (the reference's flat substring
  // automaton, src/filter.rs:1124-1207, here over the whole address with the
  // suffixes and class prefixes in the same automaton).
  static std::string dictionary_source(const FilterSet &fs, size_t lo, size_t hi) {
    static const char kSym[] = "abcdefghijklmnopqrstuvwxyz234567";
    std::string src;
    for (size_t n = lo; n < hi; ++n) {
      const uint32_t idx = fs.general[n];
      const Filter &f = fs.filters[idx];
      if (!src.empty()) src += '|';
      if (f.anchor == kStart) src += '^';
      for (uint32_t set : f.sets) {
        if (set == 0xffffffffu) {
          src += '.';
        } else if (__builtin_popcount(set) == 1) {
          src += kSym[__builtin_ctz(set)];
        } else {
          src += '[';
          for (int c = 0; c < 32; ++c)
            if ((set >> c) & 1u) src += kSym[c];
          src += ']';
        }
      }
      if (f.anchor == kEnd) src += '$';
    }
    return src;
  }

  // Expressions alone, each settled by the key and guarded by few enough
  // literal prefixes: the prefixes go to the key tables (the reference's
  // Path::Bits, src/expression.rs:96-107).  Empty string when taken.
  std::string compile_key_proxy(const FilterSet &fs) {
    if (fs.regexes.empty() || !fs.fast.empty() || !fs.general.empty()) return "no";
    if (fs.regexes.size() > (size_t)kMaxRegexes) return "no";
    std::vector<std::string> lines;
    for (uint32_t i : fs.regexes) {
      const std::vector<std::string> w = fs.filters[i].regex->literal_prefixes(4096, 1.0 / 1024);
      if (w.empty()) return "no";
      lines.insert(lines.end(), w.begin(), w.end());
    }
    if (!build_filter_set(lines, 0, proxy_set).empty() || proxy_set.fast.empty() ||
        !proxy_set.general.empty())
      return "no";
    const std::string e = compile(proxy_set);
    if (!e.empty() || (mode & kModePattern)) return "no";
    key_proxy = true;
    add_automata(fs);
    return "";
  }

  // General filters [lo, hi) as one automaton, or, when that exceeds the state
  // limit, as the automata of its two halves; at most `budget` in all.
  std::string compile_dictionary(const FilterSet &fs, size_t lo, size_t hi, size_t budget) {
    RegexLimits lim;
    lim.program = 4000000;
    lim.states = kRegexMaxStates;
    lim.proxies = false;
    std::shared_ptr<const Regex> re;
    const std::string e = Regex::compile(dictionary_source(fs, lo, hi), re, lim);
    if (e.empty()) {
      if (dictionaries.size() >= budget) return "more automata than the device holds (16)";
      dictionaries.push_back(re);
      return "";
    }
    if (hi - lo < 2 || e.find("automaton states") == std::string::npos) return e;
    const size_t mid = lo + (hi - lo) / 2;
    const std::string a = compile_dictionary(fs, lo, mid, budget);
    return a.empty() ? compile_dictionary(fs, mid, hi, budget) : a;
  }

  // Empty string on success.
  std::string compile(const FilterSet &fs) {
    if (!fs.regexes.empty() && fs.fast.empty() && fs.general.empty()) {
      DevFilterHost k;
      if (k.compile_key_proxy(fs).empty()) {
        *this = std::move(k);
        return "";
      }
    }
    if (fs.fast.size() == 1) {
      mode |= kModePrefix;
      prefix_words(fs.filters[fs.fast[0]], want, mask);
    } else if (!fs.fast.empty()) {
      mode |= kModeBitmap;
      // A map of at least 24 bits, at any occupancy: behind it the exact
      // table decides, so the map only sets how many candidates reach that
      // table.  The reference's own index is taken when it is wider (up to 32
      // bits); a narrower one (its automatic choice stops at 1% occupancy,
      // 20 bits for 10,000 prefixes) would send about 16 times as many.
      const bool theirs = fs.index_bits > 24;
      const std::vector<uint64_t> w = theirs ? fs.words : bitmap_for(fs, 24);
      bits = theirs ? fs.index_bits : 24;
      set_bitmap(w);
      if (fs.buckets.size() > (size_t)kMaxBuckets)
        return "more distinct prefix lengths than the device table holds";
      uint32_t run = 0;  // the most keys behind one directory entry
      for (const Bucket &b : fs.buckets) {
        // the directory over the top D bits of the bucket's sorted keys,
        // about one key an entry: entry t is the first key whose top D bits
        // are t or more, entry 2^D the count
        const uint32_t n = (uint32_t)b.keys.size();
        uint32_t D = 1;
        while (D < kDirMaxBits && (1u << D) < n) ++D;
        const uint32_t shift = 64 - D, at = (uint32_t)dirs.size();
        dirs.resize(at + (1u << D) + 1);
        uint32_t i = 0;
        for (uint32_t t = 0; t <= (1u << D); ++t) {
          while (i < n && (b.keys[i] >> shift) < t) ++i;
          dirs[at + t] = i;
          if (t) run = std::max(run, dirs[at + t] - dirs[at + t - 1]);
        }
        const uint32_t rec[kBucketWords] = {(uint32_t)keys.size(), n, (uint32_t)b.mask,
                                            (uint32_t)(b.mask >> 32), at, shift, 0, 0};
        buckets.insert(buckets.end(), rec, rec + kBucketWords);
        keys.insert(keys.end(), b.keys.begin(), b.keys.end());
      }
      dir_steps = 0;  // halvings that bring a run of `run` keys to one place
      while ((1u << dir_steps) <= run) ++dir_steps;
    }
    if (!fs.general.empty() || !fs.regexes.empty()) {
      if (fs.regexes.size() > (size_t)kMaxRegexes)
        return "at most 16 regex: filters per search";
      mode |= kModePattern;
      bool dict = fs.general.size() > (size_t)kMaxPatterns;
      if (!dict && !build_banks(fs, banks, &proxyless).empty()) dict = true;
      if (dict) {
        // too many for the banks: automata over the general filters
        const std::string e =
            compile_dictionary(fs, 0, fs.general.size(), kMaxRegexes - fs.regexes.size());
        if (!e.empty()) return "the general filters as automata: " + e;
        dictionary_forms = fs.general.size();
        banks.clear();
        const std::string eb = build_banks(fs, banks, &proxyless, false);
        if (!eb.empty()) return eb;
        for (const auto &d : dictionaries) add_automaton(*d);
      } else {
        for (uint32_t i : fs.general) {
          DevPattern p = dev_pattern(fs.filters[i]);
          pats.push_back(p.len);
          pats.push_back(p.first);
          pats.push_back(p.last);
          pats.push_back(p.key_last);
          for (int s = 0; s < kAddressLen; ++s) pats.push_back(p.sets[s]);
        }
      }
      add_automata(fs);
      regex_every = proxyless || !dictionaries.empty();
    }
    if (!mode) return "the filter set is empty";
    need = need_rate(kNeedSamples);
    park_all = (mode & kModePattern) && need >= kParkAllRate;
    build_window_prefilter(fs);
    return "";
  }

  // The window prefilter, for a dictionary that is only plain substrings -
  // unanchored, one symbol per position, 5 to 49 symbols - walked in global
  // memory (an automaton in LDS walks faster than this saves), with no
  // expression and no bank beside it.  Its form - window length W and
  // placement stride S, W + S - 1 at most the shortest word - is the one that
  // costs least: probes per key (a few instructions and two LDS reads each)
  // against the keys its false hits send on.
  // - The key stage (hash_key) sends a key whose window hits to the address
  //   stage.  A form is worth it while the expected share of such keys is at
  //   most kWindowParkMax; each costs about what 400 probes do.
  // - When nearly every candidate goes to the address stage anyway
  //   (park_all), the address stage uses the maps instead (hash_addr): its
  //   whole walk runs only in waves where a lane's window hits, worth it while
  //   at most half the waves do; a wave's walk costs about what 250 probes do.
  void build_window_prefilter(const FilterSet &fs) {
    hash.clear();
    hash_key = hash_addr = false;
    hash_lmax = hash_ws = 0;
    window_park = window_share = 0.0;
    if (dictionaries.empty() || !fs.regexes.empty() || !banks.empty() || regex.empty() ||
        first_states <= kLdsStates)
      return;
    uint32_t lmax = 0, lmin = 49;
    for (uint32_t i : fs.general) {
      const Filter &f = fs.filters[i];
      if (f.anchor != kAnywhere || f.sets.size() < 5 || f.sets.size() > 49) return;
      for (uint32_t set : f.sets)
        if (__builtin_popcount(set) != 1) return;
      lmax = std::max<uint32_t>(lmax, (uint32_t)f.sets.size());
      lmin = std::min<uint32_t>(lmin, (uint32_t)f.sets.size());
    }
    const uint32_t end = park_all ? 56 : 49;  // the symbols probed
    double best = 1e300, best_share = 1.0;
    uint32_t best_ws = 0;
    std::vector<uint32_t> best_maps;
    for (uint32_t W = 5; W <= 7 && W <= lmin; ++W)
      for (uint32_t S = 1; S <= 4 && W + S - 1 <= lmin; ++S) {
        std::vector<uint32_t> maps = window_maps(fs, W, S);
        const double p = window_probe_rate(maps, W);
        const uint32_t probes = (end - W) / S + 1;
        const double keys = 1.0 - std::pow(1.0 - p, (double)probes);
        const double share = park_all ? 1.0 - std::pow(1.0 - keys, 64.0) : keys;
        const double cost = probes + share * (park_all ? 250.0 : 400.0);
        if (cost < best) {
          best = cost;
          best_share = share;
          best_ws = 8 * W + S;
          best_maps.swap(maps);
        }
      }
    if (!best_ws || best_share > (park_all ? 0.5 : kWindowParkMax)) return;
    hash.swap(best_maps);
    hash_lmax = lmax;
    hash_ws = best_ws;
    window_park = park_all ? 0.0 : best_share;
    window_share = best_share;
    hash_key = !park_all;
    hash_addr = park_all;
  }

  // The two maps of the form (W, S): each word's windows at its symbols
  // 0..S-1, the first 18 of their 5W bits in one map and the last 18 in the
  // other.
  static std::vector<uint32_t> window_maps(const FilterSet &fs, uint32_t W, uint32_t S) {
    std::vector<uint32_t> maps(2 * kHashSliceWords, 0u);
    for (uint32_t i : fs.general)
      for (uint32_t off = 0; off < S; ++off) {
        uint64_t v = 0;  // W symbols from `off`, big-endian
        for (uint32_t j = 0; j < W; ++j)
          v = (v << 5) | (uint32_t)__builtin_ctz(fs.filters[i].sets[off + j]);
        const uint32_t a = (uint32_t)(v >> (5 * W - kHashSliceBits));
        const uint32_t b = (uint32_t)(v & ((1u << kHashSliceBits) - 1));
        maps[a >> 5] |= 1u << (a & 31);
        maps[kHashSliceWords + (b >> 5)] |= 1u << (b & 31);
      }
    return maps;
  }

  // The chance that a uniformly random W-symbol window hits both maps.  The
  // two slices share their middle 36 - 5W bits (1 for W = 7, 11 for W = 5),
  // so it is the sum over the first map's entries of the share of the second
  // map's entries that agree with it there.
  static double window_probe_rate(const std::vector<uint32_t> &maps, uint32_t W) {
    const uint32_t ov = 36 - 5 * W, rest = kHashSliceBits - ov;
    std::vector<uint32_t> agree((size_t)1 << ov, 0u);  // second-map entries per shared value
    for (uint32_t i = 0; i < kHashSliceWords; ++i)
      for (uint32_t m = maps[kHashSliceWords + i]; m; m &= m - 1) {
        const uint32_t b = (i << 5) | (uint32_t)__builtin_ctz(m);
        agree[b >> rest]++;
      }
    double sum = 0;
    for (uint32_t i = 0; i < kHashSliceWords; ++i)
      for (uint32_t m = maps[i]; m; m &= m - 1) {
        const uint32_t a = (i << 5) | (uint32_t)__builtin_ctz(m);
        sum += agree[a & ((1u << ov) - 1)];
      }
    return sum / (double)(1u << kHashSliceBits) / (double)(1u << rest);
  }

  // `bits` must be set first.  Slot s of the map is bit s & 31 of word s >> 5
  // (the low half of the reference's 64-bit word first); its coarse slot is
  // s >> (bits - coarse_bits), so a coarse bit is set when any full slot
  // under it is, and a clear coarse bit rules the key out.
  void set_bitmap(const std::vector<uint64_t> &w) {
    bitmap.clear();
    for (uint64_t v : w) {
      bitmap.push_back((uint32_t)v);
      bitmap.push_back((uint32_t)(v >> 32));
    }
    coarse_bits = std::min<uint32_t>(bits, kCoarseBits);
    coarse.assign(((size_t)1 << coarse_bits) / 32 + 1, 0u);
    const uint32_t shift = bits - coarse_bits;
    for (size_t i = 0; i < bitmap.size(); ++i)
      for (uint32_t m = bitmap[i]; m; m &= m - 1) {
        const uint64_t slot = ((uint64_t)i << 5) | (uint32_t)__builtin_ctz(m);
        const uint64_t c = slot >> shift;
        coarse[c >> 5] |= 1u << (c & 31);
      }
  }

  template <class T>
  static hipError_t upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    if (src.empty()) return hipSuccess;
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
  }

  hipError_t to_device(hipStream_t s) {
    hipError_t e = upload(&d_bitmap, bitmap, s);
    if (e == hipSuccess) e = upload(&d_coarse, coarse, s);
    if (e == hipSuccess) e = upload(&d_hash, hash, s);
    if (e == hipSuccess) e = upload(&d_buckets, buckets, s);
    if (e == hipSuccess) e = upload(&d_keys, keys, s);
    if (e == hipSuccess) e = upload(&d_dirs, dirs, s);
    if (e == hipSuccess) e = upload(&d_pats, pats, s);
    if (e == hipSuccess) e = upload(&d_regex, regex, s);
    if (e == hipSuccess) e = upload(&d_dfa, dfa, s);
    if (e == hipSuccess) e = upload(&d_dfa_end, dfa_end, s);
    if (e == hipSuccess) e = upload(&d_live, live, s);
    if (e == hipSuccess && !banks.empty()) {
      std::vector<uint32_t> words(banks.size() * kBankWords);
      std::memcpy(words.data(), banks.data(), words.size() * 4);
      e = upload(&d_banks, words, s);
      // the staging vector dies here: wait for the copy that reads it
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    return e;
  }

  void free_device() {
    (void)hipFree(d_bitmap);
    (void)hipFree(d_coarse);
    (void)hipFree(d_hash);
    d_coarse = d_hash = nullptr;
    (void)hipFree(d_buckets);
    (void)hipFree(d_keys);
    (void)hipFree(d_dirs);
    d_dirs = nullptr;
    (void)hipFree(d_pats);
    (void)hipFree(d_banks);
    (void)hipFree(d_regex);
    (void)hipFree(d_dfa);
    (void)hipFree(d_dfa_end);
    (void)hipFree(d_live);
    d_bitmap = d_buckets = d_pats = d_banks = d_regex = d_live = d_dfa_end = nullptr;
    d_dfa = nullptr;
    d_keys = nullptr;
  }

  DevFilterArgs args() const {
    DevFilterArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = mode;
    std::memcpy(a.want, want, sizeof want);
    std::memcpy(a.mask, mask, sizeof mask);
    a.bits = bits;
    a.bitmap = d_bitmap;
    a.coarse = d_coarse;
    a.hash = d_hash;
    a.hash_lmax = hash_lmax;
    a.hash_ws = hash_ws;
    a.hash_key = d_hash && hash_key ? 1u : 0u;
    a.hash_addr = d_hash && hash_addr ? 1u : 0u;
    a.coarse_bits = d_coarse ? coarse_bits : 0u;
    a.nbuckets = (uint32_t)(buckets.size() / kBucketWords);
    a.buckets = d_buckets;
    a.keys = d_keys;
    a.dirs = d_dirs;
    a.dir_steps = dir_steps;
    a.nbanks = (uint32_t)banks.size();
    a.banks = d_banks;
    a.npat = (uint32_t)(pats.size() / kPatWords);
    a.pats = d_pats;
    a.nregex = (uint32_t)(regex.size() / 3);
    a.regex_live = d_live;
    a.lds_states = !regex.empty() && first_states <= kLdsStates ? first_states : 0u;
    a.regex_every = regex_every ? 1u : 0u;
    a.key_proxy = key_proxy ? 1u : 0u;
    a.regex = d_regex;
    a.dfa = d_dfa;
    a.dfa_end = d_dfa_end;
    a.defer_sign = 0;
    a.park_all = park_all ? 1u : 0u;
    return a;
  }

  // The fraction of keys whose key stage leaves a checksum placement possible
  // (the lanes that park), over `samples` pseudo-random keys: the key stage of
  // the search kernels run on the host (the general forms and expressions).
  double need_rate(uint32_t samples) const {
    if (!(mode & kModePattern) || samples == 0) return 0.0;
    const uint32_t nb = (uint32_t)banks.size(), nre = (uint32_t)(regex.size() / 3);
    const uint32_t K = !regex.empty() && first_states <= kLdsStates ? first_states : 0u;
    uint64_t st = 0x243f6a8885a308d3ull;
    uint32_t need = 0;
    for (uint32_t i = 0; i < samples; ++i) {
      uint32_t kw[8];
      for (int l = 0; l < 8; ++l) {
        st += 0x9e3779b97f4a7c15ull;  // splitmix64
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        kw[l] = (uint32_t)(z ^ (z >> 31));
      }
      kw[7] &= ~0x80u;  // the sign bit of the packed key is clear in the key stage
      bool kh, nd, maybe;
      scan_key(kw, banks.data(), nb, kh, nd, maybe);
      if (nre) {
        maybe = (maybe || regex_every) && !kh;
        if (maybe) {
          const uint32_t r = regex_key(kw, regex.data(), dfa.data(), live.data(), nre,
                                       dfa.data() + regex[0], K);
          kh |= r == 2;
          nd |= r == 1;
        }
      }
      need += nd && !kh;
    }
    return (double)need / samples;
  }
};

}  // namespace og
