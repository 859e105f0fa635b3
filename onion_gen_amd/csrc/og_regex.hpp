// Regular expressions over the 56-symbol address: the reference's `regex:` form
// (src/expression.rs, src/filter.rs:353-375, 884-1005).
//
// The reference matches with the `regex` crate (1.13.1, `regex::bytes` with
// Unicode off) and proves its prefilter literals with `regex-syntax` 0.8.11's
// extractor.  Neither exists here, so this file restates what the search needs
// of them:
//
//   * a parser for the crate's syntax, with Unicode off: literals, `.`,
//     classes (ranges, negation, nested classes, `[:name:]`, `&&` `--` `~~`),
//     the Perl classes \d \s \w and their negations, groups (capturing,
//     named, non-capturing, with flags i m s U u x R), alternation, the
//     repetitions * + ? {n} {n,} {n,m} {,m} (lazy forms too: laziness does not
//     change whether a match exists), and the looks ^ $ \A \z \b \B \< \>
//     \b{start} \b{end} \b{start-half} \b{end-half}.  Look-around and
//     back-references are refused, as the crate refuses them;
//   * a Pike VM that answers `is_match` over any text (the host's matcher);
//   * a deterministic automaton over the 32 symbol values for the address
//     (the device's matcher): the address is always 56 symbols, all of them
//     word characters, so a look depends only on whether the position is 0,
//     56 or in between, and the automaton can be built ahead of time;
//   * obligatory class sequences ("proxies"): every match begins with a word
//     of one of them, at symbol 0 when the expression is anchored.  This is
//     the reference's literal extraction (src/expression.rs:55-70, 300-380)
//     with symbol sets in place of bytes, which is what the device's
//     automaton banks take anyway: `^a[2-7]{3}shop` is one 8-position
//     sequence instead of 216 literals.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

namespace og {

// One obligatory class sequence: a symbol set per position (at most 32).
struct RegexProxy {
  bool anchored = false;       // at symbol 0; otherwise anywhere
  std::vector<uint32_t> sets;
};

// Device automaton entries: bits 0..14 the next state, bit 15 set when the
// text may end in that state (the acceptance at symbol 56).  State 0 can no
// longer match, state 1 has matched (both absorbing).
constexpr uint32_t kRegexDead = 0;
constexpr uint32_t kRegexMatched = 1;
constexpr uint32_t kRegexEndBit = 0x8000u;
constexpr uint32_t kRegexMaxStates = 32767;

// Size limits of one compiled expression.  The defaults suit an expression
// typed by a user; a dictionary of general filters compiled into one
// automaton (og_devfilter.hpp) raises them and skips the proxies.
struct RegexLimits {
  size_t program = 20000;  // program instructions
  uint32_t states = 4096;  // automaton states
  bool proxies = true;
};

class Regex {
 public:
  // Empty string on success, else the message (prefixed like the reference's
  // ExpressionError: "the regular expression does not parse: ..." or "no
  // address can ever match this expression: ...").
  static std::string compile(const std::string &source, std::shared_ptr<const Regex> &out,
                             const RegexLimits &limits = RegexLimits());

  // Whether the expression matches anywhere in `text` (any bytes).
  bool find(const char *text, int n) const;

  const std::string &source() const { return source_; }
  // Every match begins at symbol 0.
  bool anchored() const { return anchored_; }
  // The 49 key-derived symbols settle every address (src/expression.rs:152-160:
  // anchored, no match longer than 49 symbols, no look but ^).
  bool key_decidable() const { return key_decidable_; }

  // The automaton: [states][32] entries, and its first state.
  const std::vector<uint16_t> &dfa() const { return dfa_; }
  uint32_t dfa_start() const { return dfa_start_; }
  uint32_t dfa_states() const { return (uint32_t)(dfa_.size() / 32); }
  // Bit 2s+b: from state s after symbol 50, a match can still be completed
  // over symbols 51..55 when the top bit of symbol 51 is b (symbols 52 and 53
  // free, 54 one of a/i/q/y, 55 'd').  The key stage asks it so that only a
  // candidate whose match can still reach the checksum pays for it.
  const std::vector<uint32_t> &tail_live() const { return live_; }

  // The proxies fitted into `budget` automaton-bank bits (shortened, then
  // merged, as the reference shortens literals past its limit), and the share
  // of random addresses they are estimated to pass.  Empty when no proxy is
  // worth its bits: the device then runs the automaton on every candidate.
  std::vector<RegexProxy> proxies(int budget) const;
  double proxy_rate() const { return proxy_rate_; }

  // The anchored proxies of a key-decidable expression expanded into literal
  // base32 prefixes, cut to the longest length at which there are at most
  // `limit` of them.  Empty when the expression is not key-decidable, has no
  // anchored proxy, or the cut prefixes would pass more than `max_rate` of
  // the addresses.  Every match begins with one of them (the reference's
  // Literals::Prefix, src/expression.rs:57-70).
  std::vector<std::string> literal_prefixes(size_t limit, double max_rate) const;

  // One line: the proxies (classes as [..]) or why there are none.
  std::string describe() const;

  // One instruction of the compiled program (a Thompson NFA).
  struct Inst {
    uint8_t op = 0;       // char, split, jump, match, assert
    uint8_t look = 0;     // assert: which look
    uint32_t mask = 0;    // char: accepted symbol values
    uint32_t x = 0, y = 0;
  };

 private:
  std::string source_;
  std::vector<Inst> prog_;
  bool anchored_ = false;
  bool key_decidable_ = false;
  std::vector<uint16_t> dfa_;
  uint32_t dfa_start_ = 0;
  std::vector<uint32_t> live_;
  std::vector<RegexProxy> proxies_;
  double proxy_rate_ = 1.0;
  std::string why_none_;
  friend struct RegexBuilder;
};

}  // namespace og
