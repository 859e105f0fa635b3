// Regular expressions for the `regex:` form: parser, Pike VM, the device
// automaton and the proxy extraction (the header gives the scope).
//
// Reference behaviour followed here:
//   Expression::parse            src/expression.rs:126-178 (minimum length past
//                                56 refused, unmatchable branches refused, the
//                                key-text haystack rule)
//   classify_literals / shorten  src/expression.rs:300-380 (empty literal: no
//                                prefilter; too many: shortened, not dropped)
//   only_safe_looks              src/expression.rs:263-279
#include "og_regex.hpp"

#include <algorithm>
#include <bitset>
#include <cctype>
#include <cstring>
#include <iterator>
#include <map>
#include <unordered_map>

namespace og {

namespace {

constexpr int kAddrLen = 56;
constexpr int kKeySymbols = 49;
constexpr size_t kMaxLits = 64;        // class sequences the extraction keeps
constexpr size_t kMaxLitLen = 32;      // one automaton bank
constexpr uint32_t kAll = 0xffffffffu;

const char kAlphabet[] = "abcdefghijklmnopqrstuvwxyz234567";

enum Op : uint8_t { kChar, kSplit, kJmp, kMatch, kAssert };
enum Look : uint8_t {
  kBegin, kEndText, kWordB, kNotWordB, kWordStart, kWordEnd, kStartHalf, kEndHalf
};
enum Kind : uint8_t { kEmpty, kSet, kLook, kConcat, kAlt, kRepeat };

struct Node {
  Kind kind = kEmpty;
  uint32_t mask = 0;  // kSet: symbol values
  Look look = kBegin;
  int min = 0, max = 0;  // kRepeat; max < 0 is unbounded
  int culprit = -1;      // kSet made from one byte no address holds
  std::vector<int> kids;
};

using Bytes = std::bitset<256>;

bool is_word_byte(int c) { return std::isalnum(c) || c == '_'; }

uint32_t mask_of(const Bytes &b) {
  uint32_t m = 0;
  for (int s = 0; s < 32; ++s)
    if (b[(unsigned char)kAlphabet[s]]) m |= 1u << s;
  return m;
}

Bytes fold(Bytes b) {
  for (int c = 'a'; c <= 'z'; ++c) {
    const int u = c - 'a' + 'A';
    if (b[c] || b[u]) b[c] = b[u] = true;
  }
  return b;
}

Bytes range(int lo, int hi) {
  Bytes b;
  for (int c = lo; c <= hi; ++c) b[c] = true;
  return b;
}

Bytes perl_class(char c) {
  Bytes b;
  switch (std::tolower((unsigned char)c)) {
    case 'd': b = range('0', '9'); break;
    case 's': for (int x : {' ', '\t', '\n', '\v', '\f', '\r'}) b[x] = true; break;
    default:  // w
      b = range('0', '9') | range('a', 'z') | range('A', 'Z');
      b['_'] = true;
  }
  return std::isupper((unsigned char)c) ? ~b : b;
}

bool posix_class(const std::string &name, Bytes &out) {
  static const std::map<std::string, Bytes> kNames = [] {
    std::map<std::string, Bytes> m;
    const Bytes digit = range('0', '9'), lower = range('a', 'z'), upper = range('A', 'Z');
    Bytes space, punct, blank, cntrl = range(0, 31), word = digit | lower | upper, xdigit;
    for (int x : {' ', '\t', '\n', '\v', '\f', '\r'}) space[x] = true;
    blank[' '] = blank['\t'] = true;
    cntrl[127] = true;
    word['_'] = true;
    xdigit = digit | range('a', 'f') | range('A', 'F');
    for (int c = 33; c < 127; ++c)
      if (!std::isalnum(c)) punct[c] = true;
    m["alnum"] = digit | lower | upper;
    m["alpha"] = lower | upper;
    m["ascii"] = range(0, 127);
    m["blank"] = blank;
    m["cntrl"] = cntrl;
    m["digit"] = digit;
    m["graph"] = range(33, 126);
    m["lower"] = lower;
    m["print"] = range(32, 126);
    m["punct"] = punct;
    m["space"] = space;
    m["upper"] = upper;
    m["word"] = word;
    m["xdigit"] = xdigit;
    return m;
  }();
  auto it = kNames.find(name);
  if (it == kNames.end()) return false;
  out = it->second;
  return true;
}

// ------------------------------------------------------------------ parser

class Parser {
 public:
  Parser(const std::string &s, std::vector<Node> &nodes) : s_(s), n_(nodes) {}

  int parse(std::string &err) {
    const int root = alt();
    if (err_.empty() && !eof()) fail(s_[i_] == ')' ? "unopened group" : "unexpected character");
    err = err_;
    return root;
  }

 private:
  const std::string &s_;
  std::vector<Node> &n_;
  size_t i_ = 0;
  std::string err_;
  bool icase_ = false, verbose_ = false;
  int depth_ = 0;

  bool eof() const { return i_ >= s_.size(); }
  char peek(size_t ahead = 0) const { return i_ + ahead < s_.size() ? s_[i_ + ahead] : '\0'; }
  void fail(const std::string &m) {
    if (err_.empty()) err_ = m + " at position " + std::to_string(i_);
  }
  int add(Node n) {
    n_.push_back(std::move(n));
    return (int)n_.size() - 1;
  }
  int empty() { return add(Node()); }
  int set_node(uint32_t mask, int culprit = -1) {
    Node n;
    n.kind = kSet;
    n.mask = mask;
    n.culprit = mask ? -1 : culprit;
    return add(n);
  }
  int look_node(Look l) {
    Node n;
    n.kind = kLook;
    n.look = l;
    return add(n);
  }
  int byte_node(int c) {
    Bytes b;
    b[c] = true;
    if (icase_) b = fold(b);
    return set_node(mask_of(b), c);
  }
  void skip_ws() {
    if (!verbose_) return;
    while (!eof()) {
      if (std::isspace((unsigned char)s_[i_])) {
        ++i_;
      } else if (s_[i_] == '#') {
        while (!eof() && s_[i_] != '\n') ++i_;
      } else {
        break;
      }
    }
  }

  int alt() {
    std::vector<int> branches{concat()};
    while (err_.empty() && !eof() && s_[i_] == '|') {
      ++i_;
      branches.push_back(concat());
    }
    if (branches.size() == 1) return branches[0];
    Node n;
    n.kind = kAlt;
    n.kids = branches;
    return add(n);
  }

  int concat() {
    Node n;
    n.kind = kConcat;
    for (;;) {
      skip_ws();
      if (!err_.empty() || eof() || s_[i_] == '|' || s_[i_] == ')') break;
      bool flags_only = false;
      const int a = atom(flags_only);
      if (!err_.empty()) break;
      if (flags_only) {
        skip_ws();
        if (!eof() && std::strchr("*+?{", s_[i_])) fail("repetition operator missing expression");
        continue;
      }
      n.kids.push_back(repeat_of(a));
    }
    if (n.kids.empty()) return empty();
    if (n.kids.size() == 1) return n.kids[0];
    return add(n);
  }

  bool number(int &out) {
    size_t start = i_;
    long v = 0;
    while (!eof() && std::isdigit((unsigned char)s_[i_])) {
      v = v * 10 + (s_[i_] - '0');
      if (v > 100000) {
        fail("repetition count too large");
        return false;
      }
      ++i_;
    }
    out = (int)v;
    return i_ > start;
  }

  int repeat_of(int a) {
    bool quantified = false;
    for (;;) {
      skip_ws();
      if (!err_.empty() || eof()) break;
      const char c = s_[i_];
      int lo = 0, hi = -1;
      if (c == '*') {
        ++i_;
      } else if (c == '+') {
        lo = 1;
        ++i_;
      } else if (c == '?') {
        hi = 1;
        ++i_;
      } else if (c == '{') {
        const size_t open = i_;
        ++i_;
        skip_ws();
        const bool have_lo = number(lo);
        skip_ws();
        if (!err_.empty()) return a;
        if (!eof() && s_[i_] == ',') {
          ++i_;
          skip_ws();
          int h = 0;
          if (number(h)) hi = h;
          else if (!have_lo) {
            i_ = open;
            fail("invalid counted repetition");
            return a;
          }
          skip_ws();
        } else {
          if (!have_lo) {
            i_ = open;
            fail("invalid counted repetition");
            return a;
          }
          hi = lo;
        }
        if (!err_.empty()) return a;
        if (eof() || s_[i_] != '}') {
          i_ = open;
          fail("unclosed counted repetition");
          return a;
        }
        ++i_;
        if (hi >= 0 && lo > hi) {
          i_ = open;
          fail("invalid repetition count range");
          return a;
        }
      } else {
        break;
      }
      if (quantified) {
        fail("repetition operator missing expression");
        return a;
      }
      quantified = true;
      if (!eof() && s_[i_] == '?') ++i_;  // lazy: the same language
      Node n;
      n.kind = kRepeat;
      n.min = lo;
      n.max = hi;
      n.kids = {a};
      a = add(n);
    }
    return a;
  }

  // A group's flags: (?flags) sets them until the group ends, (?flags:...)
  // only inside.  Returns false when it was the first form.
  bool flags(bool &scoped) {
    bool neg = false, icase = icase_, verbose = verbose_;
    for (;;) {
      if (eof()) {
        fail("unclosed group flags");
        return false;
      }
      const char c = s_[i_++];
      if (c == ')' || c == ':') {
        scoped = c == ':';
        icase_ = icase;
        verbose_ = verbose;
        return true;
      }
      if (c == '-') {
        if (neg) {
          fail("repeated negation in flags");
          return false;
        }
        neg = true;
      } else if (c == 'i') {
        icase = !neg;
      } else if (c == 'x') {
        verbose = !neg;
      } else if (!std::strchr("msUuR", c)) {
        --i_;
        fail("unrecognized flag");
        return false;
      }
    }
  }

  int group(bool &flags_only) {
    const size_t open = i_;
    ++i_;  // '('
    const bool icase = icase_, verbose = verbose_;
    if (peek() == '?') {
      ++i_;
      const char c = peek();
      if (c == '=' || c == '!' || (c == '<' && (peek(1) == '=' || peek(1) == '!'))) {
        fail("look-around, including look-ahead and look-behind, is not supported");
        return -1;
      }
      if (c == 'P' || c == '<') {
        i_ += c == 'P' ? 2 : 1;
        if (c == 'P' && s_[i_ - 1] != '<') {
          fail("invalid capture group name");
          return -1;
        }
        const size_t start = i_;
        while (!eof() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' ||
                          s_[i_] == '.' || s_[i_] == '[' || s_[i_] == ']'))
          ++i_;
        if (eof() || s_[i_] != '>' || i_ == start) {
          fail("invalid capture group name");
          return -1;
        }
        ++i_;
      } else {
        bool scoped = false;
        if (!flags(scoped)) return -1;
        if (!scoped) {
          flags_only = true;  // applies to the rest of the enclosing group
          return -1;
        }
      }
    }
    if (++depth_ > 250) {
      fail("nesting too deep");
      return -1;
    }
    const int inner = alt();
    --depth_;
    if (!err_.empty()) return -1;
    if (eof() || s_[i_] != ')') {
      i_ = open;
      fail("unclosed group");
      return -1;
    }
    ++i_;
    icase_ = icase;
    verbose_ = verbose;
    return inner;
  }

  // \x7f, \x{7f}: one byte value (a code point past 0xff matches no byte).
  int hex_escape() {
    int v = 0, digits = 0;
    if (peek() == '{') {
      ++i_;
      while (!eof() && std::isxdigit((unsigned char)s_[i_]) && digits < 8) {
        v = v * 16 + (std::isdigit((unsigned char)s_[i_]) ? s_[i_] - '0'
                                                           : std::tolower(s_[i_]) - 'a' + 10);
        ++i_;
        ++digits;
      }
      if (eof() || s_[i_] != '}' || digits == 0) {
        fail("invalid hexadecimal escape");
        return -1;
      }
      ++i_;
    } else {
      for (int k = 0; k < 2; ++k) {
        if (eof() || !std::isxdigit((unsigned char)s_[i_])) {
          fail("invalid hexadecimal escape");
          return -1;
        }
        v = v * 16 + (std::isdigit((unsigned char)s_[i_]) ? s_[i_] - '0'
                                                           : std::tolower(s_[i_]) - 'a' + 10);
        ++i_;
      }
    }
    return v > 255 ? 256 : v;
  }

  // The escapes that stand for one byte (inside a class or out).  -1: not one.
  int byte_escape(char c) {
    switch (c) {
      case 'n': return '\n';
      case 't': return '\t';
      case 'r': return '\r';
      case 'f': return '\f';
      case 'v': return '\v';
      case 'a': return 7;
      case 'x': return hex_escape();
      default:
        if (std::isalnum((unsigned char)c)) return -1;
        if ((unsigned char)c < 128 && !std::isspace((unsigned char)c) && c != '<' && c != '>')
          return (unsigned char)c;
        return -1;
    }
  }

  int escape() {
    ++i_;  // '\'
    if (eof()) {
      fail("incomplete escape sequence");
      return -1;
    }
    const char c = s_[i_++];
    switch (c) {
      case 'd': case 'D': case 's': case 'S': case 'w': case 'W': {
        Bytes b = perl_class(c);
        if (icase_) b = fold(b);
        return set_node(mask_of(b));
      }
      case 'A': return look_node(kBegin);
      case 'z': return look_node(kEndText);
      case 'B': return look_node(kNotWordB);
      case '<': return look_node(kWordStart);
      case '>': return look_node(kWordEnd);
      case 'b': {
        if (peek() == '{') {
          const size_t close = s_.find('}', i_);
          const std::string name = close == std::string::npos ? "" : s_.substr(i_ + 1, close - i_ - 1);
          Look l;
          if (name == "start") l = kWordStart;
          else if (name == "end") l = kWordEnd;
          else if (name == "start-half") l = kStartHalf;
          else if (name == "end-half") l = kEndHalf;
          else return look_node(kWordB);  // a repetition of \b follows
          i_ = close + 1;
          return look_node(l);
        }
        return look_node(kWordB);
      }
      case 'p': case 'P':
        --i_;
        fail("Unicode classes are not available with Unicode mode off");
        return -1;
      default: break;
    }
    if (c >= '0' && c <= '9') {
      --i_;
      fail("backreferences are not supported");
      return -1;
    }
    const int b = byte_escape(c);
    if (!err_.empty()) return -1;
    if (b < 0) {
      --i_;
      fail("unrecognized escape sequence");
      return -1;
    }
    if (b > 255) return set_node(0, -1);
    return byte_node(b);
  }

  // One class item that is a single byte (for ranges): -1 when it is not.
  int class_byte() {
    if (eof()) return -1;
    if (s_[i_] == '\\') {
      const char c = peek(1);
      if (std::strchr("dDsSwW", c)) return -1;
      i_ += 2;
      const int b = byte_escape(c);
      if (!err_.empty()) return -1;
      if (b < 0) {
        i_ -= 1;
        fail("unrecognized escape sequence in a class");
        return -1;
      }
      return b > 255 ? 256 : b;
    }
    return (unsigned char)s_[i_++];
  }

  Bytes class_union(bool first) {
    Bytes acc;
    for (;;) {
      if (eof()) {
        fail("unclosed character class");
        return acc;
      }
      const char c = s_[i_];
      if (c == ']' && !first) break;
      if ((c == '&' && peek(1) == '&') || (c == '-' && peek(1) == '-' && !first) ||
          (c == '~' && peek(1) == '~'))
        break;
      first = false;
      if (c == '[' && peek(1) == ':') {
        const size_t close = s_.find(":]", i_ + 2);
        if (close != std::string::npos) {
          std::string name = s_.substr(i_ + 2, close - i_ - 2);
          const bool neg = !name.empty() && name[0] == '^';
          if (neg) name = name.substr(1);
          Bytes b;
          if (posix_class(name, b)) {
            i_ = close + 2;
            acc |= neg ? ~b : b;
            continue;
          }
        }
      }
      if (c == '[') {
        acc |= class_body();
        if (!err_.empty()) return acc;
        continue;
      }
      if (c == '\\' && std::strchr("dDsSwW", peek(1))) {
        acc |= perl_class(peek(1));
        i_ += 2;
        continue;
      }
      const size_t at = i_;
      const int lo = class_byte();
      if (!err_.empty()) return acc;
      if (peek() == '-' && peek(1) != ']' && peek(1) != '\0' && peek(1) != '-') {
        ++i_;
        const int hi = class_byte();
        if (!err_.empty()) return acc;
        if (hi < 0 || lo > hi) {
          i_ = at;
          fail("invalid character class range");
          return acc;
        }
        if (lo <= 255) acc |= range(lo, std::min(hi, 255));
      } else if (lo >= 0 && lo <= 255) {
        acc[lo] = true;
      }
    }
    return acc;
  }

  // After '[': the whole class, negation and set operations applied.
  Bytes class_body() {
    const size_t open = i_;
    ++i_;
    bool neg = false;
    if (peek() == '^') {
      neg = true;
      ++i_;
    }
    Bytes acc = class_union(true);
    if (icase_) acc = fold(acc);
    while (err_.empty() && !eof() && s_[i_] != ']') {
      const char op = s_[i_];
      i_ += 2;
      Bytes rhs = class_union(false);
      if (icase_) rhs = fold(rhs);
      if (op == '&') acc &= rhs;
      else if (op == '-') acc &= ~rhs;
      else acc ^= rhs;
    }
    if (!err_.empty()) return acc;
    if (eof()) {
      i_ = open;
      fail("unclosed character class");
      return acc;
    }
    ++i_;  // ']'
    return neg ? ~acc : acc;
  }

  int atom(bool &flags_only) {
    const char c = s_[i_];
    switch (c) {
      case '(': return group(flags_only);
      case '[': {
        const Bytes b = class_body();
        return err_.empty() ? set_node(mask_of(b)) : -1;
      }
      case '.': ++i_; return set_node(kAll);
      case '^': ++i_; return look_node(kBegin);
      case '$': ++i_; return look_node(kEndText);
      case '\\': return escape();
      case '*': case '+': case '?': case '{':
        fail("repetition operator missing expression");
        return -1;
      default:
        ++i_;
        return byte_node((unsigned char)c);
    }
  }
};

// --------------------------------------------------------------- analysis

struct Facts {
  long min_len = 0;   // capped
  bool dead = false;  // no string of address symbols matches
  int culprit = -1;   // for a dead node: a byte every branch needs
};

Facts facts(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  Facts f;
  switch (x.kind) {
    case kEmpty: case kLook: return f;
    case kSet:
      f.min_len = 1;
      f.dead = x.mask == 0;
      f.culprit = x.culprit;
      return f;
    case kConcat:
      for (int k : x.kids) {
        const Facts g = facts(n, k);
        f.min_len = std::min(f.min_len + g.min_len, 1L << 30);
        if (g.dead && !f.dead) {
          f.dead = true;
          f.culprit = g.culprit;
        }
      }
      return f;
    case kAlt: {
      f.min_len = 1L << 30;
      f.dead = true;
      for (int k : x.kids) {
        const Facts g = facts(n, k);
        f.min_len = std::min(f.min_len, g.min_len);
        if (!g.dead) f.dead = false;
        if (g.dead && f.culprit < 0) f.culprit = g.culprit;
      }
      if (!f.dead) f.culprit = -1;
      return f;
    }
    case kRepeat: {
      const Facts g = facts(n, x.kids[0]);
      f.min_len = std::min(g.min_len * x.min, 1L << 30);
      f.dead = x.min > 0 && g.dead;
      f.culprit = f.dead ? g.culprit : -1;
      return f;
    }
  }
  return f;
}

bool zero_width(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  if (x.kind == kEmpty || x.kind == kLook) return true;
  if (x.kind == kConcat || x.kind == kAlt) {
    for (int k : x.kids)
      if (!zero_width(n, k)) return false;
    return true;
  }
  if (x.kind == kRepeat) return x.max == 0 || zero_width(n, x.kids[0]);
  return false;
}

// Every match passes ^ (or \A) before it reads a symbol.
bool starts_anchored(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  switch (x.kind) {
    case kLook: return x.look == kBegin;
    case kConcat:
      for (int k : x.kids) {
        if (starts_anchored(n, k)) return true;
        if (!zero_width(n, k)) return false;
      }
      return false;
    case kAlt:
      for (int k : x.kids)
        if (!starts_anchored(n, k)) return false;
      return !x.kids.empty();
    case kRepeat: return x.min > 0 && starts_anchored(n, x.kids[0]);
    default: return false;
  }
}

// The longest match, or -1 when unbounded (capped at 2^20).
long max_len(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  switch (x.kind) {
    case kEmpty: case kLook: return 0;
    case kSet: return x.mask ? 1 : 0;
    case kConcat: {
      long t = 0;
      for (int k : x.kids) {
        const long m = max_len(n, k);
        if (m < 0) return -1;
        t = std::min(t + m, 1L << 20);
      }
      return t;
    }
    case kAlt: {
      long t = 0;
      for (int k : x.kids) {
        const long m = max_len(n, k);
        if (m < 0) return -1;
        t = std::max(t, m);
      }
      return t;
    }
    case kRepeat: {
      const long m = max_len(n, x.kids[0]);
      if (m == 0) return 0;
      if (m < 0 || x.max < 0) return -1;
      return std::min(m * x.max, 1L << 20);
    }
  }
  return -1;
}

bool only_begin_looks(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  if (x.kind == kLook) return x.look == kBegin;
  for (int k : x.kids)
    if (!only_begin_looks(n, k)) return false;
  return true;
}

// --------------------------------------------------------------- proxies

struct Lit {
  std::vector<uint32_t> sets;
  bool exact = true;
  bool operator<(const Lit &o) const {
    return sets != o.sets ? sets < o.sets : exact < o.exact;
  }
  bool operator==(const Lit &o) const { return sets == o.sets && exact == o.exact; }
};

// The extraction's result: `known` false when nothing could be proven.
struct Seq {
  bool known = true;
  std::vector<Lit> lits;
};

void tidy(std::vector<Lit> &v) {
  for (Lit &l : v) {
    if (l.sets.size() > kMaxLitLen) {
      l.sets.resize(kMaxLitLen);
      l.exact = false;
    }
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// Too many sequences become one: the position-wise union over the shortest
// length.  Every match still begins with a word of it, and it stays exact only
// when every sequence was exact and of that length.
void merge(std::vector<Lit> &v) {
  if (v.size() <= kMaxLits) return;
  size_t len = v[0].sets.size();
  for (const Lit &l : v) len = std::min(len, l.sets.size());
  Lit m;
  m.sets.assign(len, 0);
  for (const Lit &l : v) {
    for (size_t j = 0; j < len; ++j) m.sets[j] |= l.sets[j];
    m.exact = m.exact && l.exact && l.sets.size() == len;
  }
  v = {m};
}

Seq cross(Seq a, const Seq &b) {
  if (!b.known) {
    for (Lit &l : a.lits) l.exact = false;
    return a;
  }
  size_t n = 0;
  for (const Lit &l : a.lits) n += l.exact ? b.lits.size() : 1;
  if (n > kMaxLits) {
    // stop extending: what was collected is still an obligatory prefix
    for (Lit &l : a.lits) l.exact = false;
    return a;
  }
  Seq r;
  for (const Lit &l : a.lits) {
    if (!l.exact) {
      r.lits.push_back(l);
      continue;
    }
    for (const Lit &m : b.lits) {
      Lit c = l;
      c.sets.insert(c.sets.end(), m.sets.begin(), m.sets.end());
      c.exact = m.exact;
      r.lits.push_back(std::move(c));
    }
  }
  tidy(r.lits);
  return r;
}

bool all_inexact(const Seq &s) {
  for (const Lit &l : s.lits)
    if (l.exact) return false;
  return true;
}

Seq extract(const std::vector<Node> &n, int i);

Seq extract_concat(const std::vector<Node> &n, const std::vector<int> &kids, size_t from) {
  Seq acc;
  acc.lits.push_back(Lit());
  for (size_t k = from; k < kids.size() && !all_inexact(acc); ++k) acc = cross(acc, extract(n, kids[k]));
  return acc;
}

Seq extract(const std::vector<Node> &n, int i) {
  const Node &x = n[i];
  Seq s;
  switch (x.kind) {
    case kEmpty: case kLook:
      s.lits.push_back(Lit());
      return s;
    case kSet:
      if (x.mask) s.lits.push_back(Lit{{x.mask}, true});
      return s;  // an empty set matches nothing: no sequence
    case kConcat: return extract_concat(n, x.kids, 0);
    case kAlt:
      for (int k : x.kids) {
        const Seq g = extract(n, k);
        if (!g.known) return Seq{false, {}};
        s.lits.insert(s.lits.end(), g.lits.begin(), g.lits.end());
      }
      tidy(s.lits);
      merge(s.lits);
      return s;
    case kRepeat: {
      const Seq g = extract(n, x.kids[0]);
      if (!g.known) return Seq{false, {}};
      if (x.max == 0) {
        s.lits.push_back(Lit());
        return s;
      }
      if (x.min == 0) {
        s.lits.push_back(Lit());
        for (Lit l : g.lits) {
          l.exact = l.exact && x.max == 1;
          s.lits.push_back(std::move(l));
        }
        tidy(s.lits);
        merge(s.lits);
        return s;
      }
      s = g;
      for (int r = 1; r < x.min && !all_inexact(s); ++r) s = cross(s, g);
      if (x.max != x.min)
        for (Lit &l : s.lits) l.exact = false;
      tidy(s.lits);
      return s;
    }
  }
  return Seq{false, {}};
}

double lit_prob(const std::vector<uint32_t> &sets) {
  double p = 1.0;
  for (uint32_t m : sets) p *= __builtin_popcount(m) / 32.0;
  return p;
}

// Leading and trailing wildcards constrain nothing a proxy can use (leading
// ones only matter where the position is fixed).
void trim(std::vector<uint32_t> &sets, bool anchored) {
  while (!sets.empty() && sets.back() == kAll) sets.pop_back();
  if (!anchored)
    while (!sets.empty() && sets.front() == kAll) sets.erase(sets.begin());
}

double rate_of(const std::vector<RegexProxy> &ps) {
  if (ps.empty()) return 1.0;
  double r = 0.0;
  for (const RegexProxy &p : ps) {
    if (p.sets.empty()) return 1.0;
    const double places = p.anchored ? 1.0 : (double)(kAddrLen + 1 - (int)p.sets.size());
    r += std::max(places, 1.0) * lit_prob(p.sets);
  }
  return std::min(r, 1.0);
}

// q covers p when every word of p begins with a word of q: p's occurrences
// are q's, so p adds nothing to the prefilter.
bool covers(const RegexProxy &q, const RegexProxy &p) {
  if (q.sets.size() > p.sets.size()) return false;
  for (size_t i = 0; i < q.sets.size(); ++i)
    if (p.sets[i] & ~q.sets[i]) return false;
  return true;
}

std::vector<RegexProxy> absorb(std::vector<RegexProxy> ps) {
  std::stable_sort(ps.begin(), ps.end(), [](const RegexProxy &a, const RegexProxy &b) {
    return a.sets.size() != b.sets.size() ? a.sets.size() < b.sets.size() : a.sets < b.sets;
  });
  std::vector<RegexProxy> kept;
  for (RegexProxy &p : ps) {
    bool dup = false;
    for (const RegexProxy &q : kept) dup = dup || covers(q, p);
    if (!dup) kept.push_back(std::move(p));
  }
  return kept;
}

std::vector<RegexProxy> to_proxies(const Seq &s, bool anchored) {
  std::vector<RegexProxy> out;
  if (!s.known || s.lits.empty()) return out;
  for (const Lit &l : s.lits) {
    RegexProxy p;
    p.anchored = anchored;
    p.sets = l.sets;
    trim(p.sets, anchored);
    if (p.sets.empty()) return {};  // a branch that proves nothing
    out.push_back(std::move(p));
  }
  return absorb(std::move(out));
}

int bits_of(const std::vector<RegexProxy> &ps) {
  int b = 0;
  for (const RegexProxy &p : ps) b += (int)std::min(p.sets.size(), kMaxLitLen);
  return b;
}

std::string set_text(uint32_t m) {
  if (m == kAll) return ".";
  if (__builtin_popcount(m) == 1) return std::string(1, kAlphabet[__builtin_ctz(m)]);
  std::string t = "[";
  for (int s = 0; s < 32; ++s)
    if ((m >> s) & 1u) t += kAlphabet[s];
  return t + "]";
}

}  // namespace

// ------------------------------------------------------------------ builder

struct RegexBuilder {
  std::vector<Node> nodes;
  std::vector<Regex::Inst> &prog;
  std::string err;
  size_t max_program;

  RegexBuilder(std::vector<Regex::Inst> &p, size_t max) : prog(p), max_program(max) {}

  size_t push(uint8_t op) {
    prog.push_back(Regex::Inst());
    prog.back().op = op;
    return prog.size() - 1;
  }

  void emit(int i) {
    if (!err.empty()) return;
    if (prog.size() > max_program) {
      err = "the expression is too large (its program exceeds " + std::to_string(max_program) +
            " instructions)";
      return;
    }
    const Node &x = nodes[i];
    switch (x.kind) {
      case kEmpty: return;
      case kSet: prog[push(kChar)].mask = x.mask; return;
      case kLook: prog[push(kAssert)].look = x.look; return;
      case kConcat:
        for (int k : x.kids) emit(k);
        return;
      case kAlt: {
        std::vector<size_t> jumps;
        for (size_t k = 0; k + 1 < x.kids.size(); ++k) {
          const size_t sp = push(kSplit);
          prog[sp].x = (uint32_t)sp + 1;
          emit(x.kids[k]);
          jumps.push_back(push(kJmp));
          prog[sp].y = (uint32_t)prog.size();
        }
        emit(x.kids.back());
        for (size_t j : jumps) prog[j].x = (uint32_t)prog.size();
        return;
      }
      case kRepeat: {
        const int kid = x.kids[0];
        for (int r = 0; r < x.min && err.empty(); ++r) emit(kid);
        if (x.max < 0) {
          const size_t sp = push(kSplit);
          prog[sp].x = (uint32_t)sp + 1;
          emit(kid);
          prog[push(kJmp)].x = (uint32_t)sp;
          prog[sp].y = (uint32_t)prog.size();
        } else {
          std::vector<size_t> exits;
          for (int r = x.min; r < x.max && err.empty(); ++r) {
            const size_t sp = push(kSplit);
            prog[sp].x = (uint32_t)sp + 1;
            exits.push_back(sp);
            emit(kid);
          }
          for (size_t e : exits) prog[e].y = (uint32_t)prog.size();
        }
        return;
      }
    }
  }
};

namespace {

// Closure of a thread set over the program at one position of an address:
// the kChar instructions reached, and whether kMatch was.
struct Closer {
  const std::vector<Regex::Inst> &prog;
  std::vector<uint32_t> mark;
  uint32_t gen = 0;
  std::vector<uint32_t> stack;

  explicit Closer(const std::vector<Regex::Inst> &p) : prog(p), mark(p.size() + 1, 0) {}

  // prev/next: whether the symbols before and after the position are word
  // characters; begin/end: the position is 0 / the text's length.
  bool run(const std::vector<uint32_t> &seeds, bool begin, bool end, bool prev, bool next,
           std::vector<uint32_t> &chars) {
    ++gen;
    chars.clear();
    bool match = false;
    stack.assign(seeds.rbegin(), seeds.rend());
    while (!stack.empty()) {
      const uint32_t pc = stack.back();
      stack.pop_back();
      if (pc >= prog.size()) {
        match = true;  // past the end: the program accepted
        continue;
      }
      if (mark[pc] == gen) continue;
      mark[pc] = gen;
      const Regex::Inst &in = prog[pc];
      switch (in.op) {
        case kChar: chars.push_back(pc); break;
        case kMatch: match = true; break;
        case kJmp: stack.push_back(in.x); break;
        case kSplit:
          stack.push_back(in.y);
          stack.push_back(in.x);
          break;
        case kAssert: {
          bool ok = false;
          switch (in.look) {
            case kBegin: ok = begin; break;
            case kEndText: ok = end; break;
            case kWordB: ok = prev != next; break;
            case kNotWordB: ok = prev == next; break;
            case kWordStart: ok = !prev && next; break;
            case kWordEnd: ok = prev && !next; break;
            case kStartHalf: ok = !prev; break;
            case kEndHalf: ok = !next; break;
          }
          if (ok) stack.push_back(pc + 1);
          break;
        }
      }
    }
    return match;
  }
};

}  // namespace

bool Regex::find(const char *text, int n) const {
  Closer c(prog_);
  std::vector<uint32_t> pending, chars;
  auto word = [&](int p) { return p >= 0 && p < n && is_word_byte((unsigned char)text[p]); };
  for (int pos = 0;; ++pos) {
    pending.push_back(0);  // a match may start here
    if (c.run(pending, pos == 0, pos == n, word(pos - 1), word(pos), chars)) return true;
    if (pos == n) return false;
    pending.clear();
    const char *hit = std::strchr(kAlphabet, text[pos]);
    if (!hit || !text[pos]) continue;
    const uint32_t bit = 1u << (uint32_t)(hit - kAlphabet);
    for (uint32_t pc : chars)
      if (prog_[pc].mask & bit) pending.push_back(pc + 1);
  }
}

namespace {

// Hash of a sorted thread set, for interning automaton states.
struct SetHash {
  size_t operator()(const std::vector<uint32_t> &v) const {
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint32_t x : v) h = (h ^ x) * 0x100000001b3ull;
    return (size_t)(h ^ (h >> 29));
  }
};

// The device automaton (see the header).  Raw states are the sets of threads
// pending at a position; state 0 of the raw numbering is the one at symbol 0,
// the only position where ^ holds.  A match may start at every position, so
// every state's closure includes the program start's; that part is the same
// for every state past symbol 0 and is computed once (for a dictionary of
// thousands of filters it is nearly all of the closure).
bool build_dfa(const std::vector<Regex::Inst> &prog, uint32_t max_states,
               std::vector<uint16_t> &table, uint32_t &start, std::string &err) {
  Closer c(prog);
  std::unordered_map<std::vector<uint32_t>, uint32_t, SetHash> ids;
  std::vector<std::vector<uint32_t>> pend;
  std::vector<uint8_t> mid, end;  // a match ends here (not the last position) / at symbol 56
  std::vector<std::vector<uint32_t>> next;
  std::vector<uint32_t> chars;
  const std::vector<uint32_t> zero{0};

  // all symbols are word characters: prev holds past 0, next before 56
  const bool start_mid = c.run(zero, false, false, true, true, chars);
  std::vector<uint32_t> start_targets[32];
  for (uint32_t pc : chars)
    for (uint32_t m = prog[pc].mask; m; m &= m - 1) start_targets[__builtin_ctz(m)].push_back(pc + 1);
  for (auto &t : start_targets) {
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end()), t.end());
  }
  const bool start_end = c.run(zero, false, true, true, false, chars);

  auto intern = [&](const std::vector<uint32_t> &p) -> uint32_t {
    auto it = ids.find(p);
    if (it != ids.end()) return it->second;
    const uint32_t id = (uint32_t)pend.size();
    ids.emplace(p, id);
    pend.push_back(p);
    mid.push_back(start_mid || c.run(p, false, false, true, true, chars));
    end.push_back(start_end || c.run(p, false, true, true, false, chars));
    next.emplace_back();
    return id;
  };
  // the state at symbol 0 is not interned: its empty thread set also names
  // the state "nothing pending" at later symbols
  pend.push_back({});
  mid.push_back(c.run(zero, true, false, false, true, chars));
  end.push_back(0);
  next.emplace_back();
  if (mid[0]) {
    // the expression matches the empty text at symbol 0: every address
    table.assign(2 * 32, 0);
    for (int s = 0; s < 32; ++s) table[32 + s] = (uint16_t)(kRegexMatched | kRegexEndBit);
    start = kRegexMatched;
    return true;
  }
  std::vector<uint32_t> own[32], merged;
  for (uint32_t u = 0; u < pend.size(); ++u) {
    if (u > 0 && mid[u]) continue;  // matched: absorbing, not expanded
    // the numbered table holds the raw states plus the two absorbing ones,
    // and a state number must fit the entries' 15 bits; an expansion adds at
    // most 32 states
    if (pend.size() + 32 + 2 > max_states) {
      err = "the expression needs more than " + std::to_string(max_states) + " automaton states";
      return false;
    }
    for (auto &t : own) t.clear();
    if (u == 0) {
      c.run(zero, true, false, false, true, chars);
    } else {
      c.run(pend[u], false, false, true, true, chars);
    }
    for (uint32_t pc : chars)
      for (uint32_t m = prog[pc].mask; m; m &= m - 1) own[__builtin_ctz(m)].push_back(pc + 1);
    next[u].resize(32);
    for (int s = 0; s < 32; ++s) {
      std::sort(own[s].begin(), own[s].end());
      merged.clear();
      if (u == 0) {
        merged = own[s];
      } else {
        std::set_union(own[s].begin(), own[s].end(), start_targets[s].begin(),
                       start_targets[s].end(), std::back_inserter(merged));
      }
      merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
      const uint32_t v = intern(merged);  // may grow `next`
      next[u][s] = v;
    }
  }
  // live: a state from which a match can still be completed
  const uint32_t n = (uint32_t)pend.size();
  std::vector<std::vector<uint32_t>> rev(n);
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t v : next[u]) rev[v].push_back(u);
  std::vector<uint8_t> live(n, 0);
  std::vector<uint32_t> work;
  for (uint32_t u = 0; u < n; ++u)
    if ((u > 0 && mid[u]) || end[u]) {
      live[u] = 1;
      work.push_back(u);
    }
  while (!work.empty()) {
    const uint32_t v = work.back();
    work.pop_back();
    for (uint32_t u : rev[v])
      if (!live[u]) {
        live[u] = 1;
        work.push_back(u);
      }
  }
  std::vector<uint32_t> num(n, kRegexDead);
  uint32_t count = 2;
  for (uint32_t u = 0; u < n; ++u) {
    if (u > 0 && mid[u]) num[u] = kRegexMatched;
    else if (live[u]) num[u] = count++;
  }
  table.assign((size_t)count * 32, 0);
  for (int s = 0; s < 32; ++s) table[32 + s] = (uint16_t)(kRegexMatched | kRegexEndBit);
  for (uint32_t u = 0; u < n; ++u) {
    if (num[u] < 2) continue;
    for (int s = 0; s < 32; ++s) {
      const uint32_t v = next[u][s];
      table[(size_t)num[u] * 32 + s] = (uint16_t)(num[v] | (end[v] ? kRegexEndBit : 0u));
    }
  }
  start = num[0];
  return true;
}

// The tail-liveness bits of the header, by dynamic programming backwards over
// symbols 55..51 of the numbered automaton.
std::vector<uint32_t> tail_liveness(const std::vector<uint16_t> &t) {
  const uint32_t n = (uint32_t)(t.size() / 32);
  const uint32_t aiqy = 0x01010101u, d = 1u << 3;
  std::vector<uint8_t> acc(n), nxt(n);
  // after symbol 54: symbol 55 is 'd', and the end decides
  for (uint32_t u = 0; u < n; ++u) acc[u] = (t[(size_t)u * 32 + 3] & kRegexEndBit) != 0;
  (void)d;
  auto step = [&](uint32_t set) {
    for (uint32_t u = 0; u < n; ++u) {
      uint8_t a = 0;
      for (uint32_t m = set; m && !a; m &= m - 1) {
        const uint32_t v = t[(size_t)u * 32 + __builtin_ctz(m)] & 0x7fffu;
        a = v == kRegexMatched || acc[v];
      }
      nxt[u] = a;
    }
    acc.swap(nxt);
  };
  step(aiqy);         // symbol 54
  step(0xffffffffu);  // symbol 53
  step(0xffffffffu);  // symbol 52
  std::vector<uint32_t> live((2 * (size_t)n + 31) / 32, 0);
  for (uint32_t b = 0; b < 2; ++b) {
    const uint32_t set = b ? 0xffff0000u : 0x0000ffffu;  // symbol 51 by its top bit
    for (uint32_t u = 0; u < n; ++u) {
      bool a = false;
      for (uint32_t m = set; m && !a; m &= m - 1) {
        const uint32_t v = t[(size_t)u * 32 + __builtin_ctz(m)] & 0x7fffu;
        a = v == kRegexMatched || acc[v];
      }
      if (a) live[(2 * u + b) >> 5] |= 1u << ((2 * u + b) & 31);
    }
  }
  return live;
}

}  // namespace

std::string Regex::compile(const std::string &source, std::shared_ptr<const Regex> &out,
                           const RegexLimits &limits) {
  auto r = std::make_shared<Regex>();
  r->source_ = source;
  RegexBuilder b(r->prog_, limits.program);
  std::string perr;
  Parser p(source, b.nodes);
  const int root = p.parse(perr);
  if (!perr.empty()) return "the regular expression does not parse: " + perr;
  const std::string unreachable = "no address can ever match this expression: ";
  const Facts f = facts(b.nodes, root);
  if (f.min_len > kAddrLen)
    return unreachable + "it cannot match fewer than " + std::to_string(f.min_len) +
           " symbols, and an address has " + std::to_string(kAddrLen);
  if (f.dead) {
    if (f.culprit >= 0) {
      std::string sym = f.culprit >= 32 && f.culprit < 127 ? std::string(1, (char)f.culprit)
                                                            : "\\x" + std::to_string(f.culprit);
      return unreachable + "every branch requires the symbol '" + sym +
             "', which base32 never produces";
    }
    return unreachable + "it accepts no string of base32 symbols";
  }
  b.emit(root);
  if (!b.err.empty()) return "the regular expression does not parse: " + b.err;
  b.push(kMatch);
  r->anchored_ = starts_anchored(b.nodes, root);
  const long longest = max_len(b.nodes, root);
  r->key_decidable_ = r->anchored_ && only_begin_looks(b.nodes, root) && longest >= 0 &&
                      longest <= kKeySymbols;
  std::string derr;
  if (!build_dfa(r->prog_, std::min(limits.states, kRegexMaxStates), r->dfa_, r->dfa_start_, derr))
    return "the expression is too complex for the device: " + derr;
  if (r->dfa_start_ == kRegexDead) return unreachable + "it accepts no address";
  r->live_ = tail_liveness(r->dfa_);
  if (!limits.proxies) {
    r->why_none_ = "not extracted";
    out = r;
    return "";
  }

  // proxies: the whole expression at symbol 0 when it is anchored, and any
  // suffix of its top-level sequence anywhere (a match of a.b contains a match
  // of b); the candidate passing the fewest addresses wins
  std::vector<RegexProxy> best;
  double best_rate = 1.0;
  auto consider = [&](std::vector<RegexProxy> c) {
    const double rate = rate_of(c);
    if (!c.empty() && (best.empty() || rate < best_rate ||
                       (rate == best_rate && bits_of(c) < bits_of(best)))) {
      best = std::move(c);
      best_rate = rate;
    }
  };
  const Node &top = b.nodes[root];
  if (r->anchored_) consider(to_proxies(extract(b.nodes, root), true));
  if (top.kind == kConcat) {
    for (size_t from = r->anchored_ ? 1 : 0; from < top.kids.size(); ++from)
      consider(to_proxies(extract_concat(b.nodes, top.kids, from), false));
  } else if (!r->anchored_) {
    consider(to_proxies(extract(b.nodes, root), false));
  }
  if (best.empty() || best_rate > 0.5) {
    r->why_none_ = best.empty() ? "no symbol sequence is common to every match"
                                : "the obligatory sequences pass most addresses";
    best.clear();
    best_rate = 1.0;
  }
  r->proxies_ = std::move(best);
  r->proxy_rate_ = best_rate;
  out = r;
  return "";
}

std::vector<RegexProxy> Regex::proxies(int budget) const {
  std::vector<RegexProxy> ps = proxies_;
  if (ps.empty() || bits_of(ps) <= budget) return ps;
  size_t longest = 0;
  for (const RegexProxy &p : ps) longest = std::max(longest, p.sets.size());
  // shorter obligatory prefixes are still obligatory (src/expression.rs:345-370)
  for (size_t keep = std::min(longest, kMaxLitLen); keep-- > 1;) {
    std::vector<RegexProxy> cut;
    for (RegexProxy p : ps) {
      if (p.sets.size() > keep) p.sets.resize(keep);
      trim(p.sets, p.anchored);
      if (p.sets.empty()) return {};
      cut.push_back(std::move(p));
    }
    cut = absorb(std::move(cut));
    if (bits_of(cut) <= budget) return cut;
  }
  if (budget < 1) return {};
  RegexProxy one;
  one.anchored = ps[0].anchored;
  one.sets = {0};
  for (const RegexProxy &p : ps) one.sets[0] |= p.sets[0];
  return {one};
}

std::vector<std::string> Regex::literal_prefixes(size_t limit, double max_rate) const {
  if (!key_decidable_ || proxies_.empty()) return {};
  size_t longest = 0;
  for (const RegexProxy &p : proxies_) {
    if (!p.anchored) return {};
    longest = std::max(longest, p.sets.size());
  }
  for (size_t len = longest; len > 0; --len) {
    double count = 0, rate = 0;
    for (const RegexProxy &p : proxies_) {
      double c = 1, q = 1;
      for (size_t i = 0; i < std::min(len, p.sets.size()); ++i) {
        c *= __builtin_popcount(p.sets[i]);
        q *= __builtin_popcount(p.sets[i]) / 32.0;
      }
      count += c;
      rate += q;
    }
    if (count > (double)limit) continue;
    if (rate > max_rate) return {};
    std::vector<std::string> out;
    for (const RegexProxy &p : proxies_) {
      std::vector<std::string> words{""};
      for (size_t i = 0; i < std::min(len, p.sets.size()); ++i) {
        std::vector<std::string> next;
        for (const std::string &w : words)
          for (uint32_t m = p.sets[i]; m; m &= m - 1) next.push_back(w + kAlphabet[__builtin_ctz(m)]);
        words.swap(next);
      }
      out.insert(out.end(), words.begin(), words.end());
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }
  return {};
}

std::string Regex::describe() const {
  if (proxies_.empty()) return "regex:" + source_ + ": no proxy (" + why_none_ + "), automaton on every candidate";
  std::string t = "regex:" + source_ + ":";
  for (size_t i = 0; i < proxies_.size(); ++i) {
    t += i ? " | " : " ";
    t += proxies_[i].anchored ? "^" : "";
    for (uint32_t m : proxies_[i].sets) t += set_text(m);
  }
  return t + ", " + std::to_string(dfa_states()) + " automaton states";
}

}  // namespace og
