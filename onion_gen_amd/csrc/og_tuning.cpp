// The tuning cache (src/tuning.rs): measured shapes written to
// $XDG_CACHE_HOME/onion-gen/tuning.yaml (else ~/.cache/onion-gen/tuning.yaml)
// in the reference's format
//
//   format: 1
//   entries:
//     <key>:
//       first: <u32>      offsets per work item: pairs per window (half)
//       second: <u32>     chains in flight (threads)
//       rate: <f64>       candidates per second measured
//       at: <i64>         unix seconds
//
// extended with `rounds` (windows per chain per launch), `radix` (the field
// representation the bench chose) and `slots` (the slot engine's unit slots),
// fields a reader of the reference format ignores.  Every operation is best effort: an unreadable or foreign file is
// "no entry" (src/tuning.rs:1-18).  OG_TUNING_PATH overrides the location.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "../../include/onion_gen_hip.h"

namespace {

std::string cache_path() {
  if (const char *p = std::getenv("OG_TUNING_PATH")) return p;
  std::string dir;
  if (const char *x = std::getenv("XDG_CACHE_HOME")) dir = x;
  else if (const char *h = std::getenv("HOME")) dir = std::string(h) + "/.cache";
  else return "";
  return dir + "/onion-gen/tuning.yaml";
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

std::string unquote(const std::string &s) {
  if (s.size() >= 2 && (s.front() == '"' || s.front() == '\'') && s.back() == s.front())
    return s.substr(1, s.size() - 2);
  return s;
}

// Reads the file; false when it is absent, unreadable or not format 1.
bool read_file(std::map<std::string, og_tuning_shape> &out) {
  std::string path = cache_path();
  if (path.empty()) return false;
  std::ifstream f(path);
  if (!f) return false;
  std::string line, current;
  bool format_ok = false, in_entries = false;
  while (std::getline(f, line)) {
    if (trim(line).empty() || trim(line)[0] == '#') continue;
    size_t indent = line.find_first_not_of(' ');
    std::string t = trim(line);
    size_t colon = t.rfind(':');
    if (indent == 0) {
      in_entries = false;
      if (t.rfind("format:", 0) == 0) format_ok = trim(t.substr(7)) == "1";
      else if (t == "entries:") in_entries = true;
      continue;
    }
    if (!in_entries || colon == std::string::npos) continue;
    if (indent == 2 && colon == t.size() - 1) {
      current = unquote(trim(t.substr(0, colon)));
      out[current] = og_tuning_shape{};
      continue;
    }
    if (indent >= 4 && !current.empty()) {
      size_t c = t.find(':');
      std::string k = trim(t.substr(0, c)), v = trim(t.substr(c + 1));
      og_tuning_shape &s = out[current];
      if (k == "first") s.first = (uint32_t)std::strtoul(v.c_str(), nullptr, 10);
      else if (k == "second") s.second = (uint32_t)std::strtoul(v.c_str(), nullptr, 10);
      else if (k == "rate") s.rate = std::strtod(v.c_str(), nullptr);
      else if (k == "at") s.at = std::strtoll(v.c_str(), nullptr, 10);
      else if (k == "rounds") s.rounds = (uint32_t)std::strtoul(v.c_str(), nullptr, 10);
      else if (k == "radix") s.radix = (uint32_t)std::strtoul(v.c_str(), nullptr, 10);
      else if (k == "slots") s.slots = (uint32_t)std::strtoul(v.c_str(), nullptr, 10);
    }
  }
  return format_ok;
}

void mkdirs(const std::string &path) {
  for (size_t i = 1; i < path.size(); ++i)
    if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

// The version the cache keys end in: the reference crate's (Cargo.toml:3,
// CARGO_PKG_VERSION), which a Rust host that links this library replaces with
// its own through og_tuning_set_version.
std::string g_version = "1.0.2";

}  // namespace

extern "C" {

int og_tuning_set_version(const char *version) {
  if (!version || !*version || std::strlen(version) > 64 || std::strchr(version, '/') ||
      std::strchr(version, '\n'))
    return -1;
  g_version = version;
  return 0;
}

size_t og_tuning_device_key(const char *name, const char *api, uint32_t threads, char *out,
                            size_t cap) {
  std::string safe = name ? name : "";
  for (char &c : safe)
    if (c == '/' || c == '\n') c = '_';
  const std::string key = "device/" + safe + "/" + (api ? api : "") + "/" +
                          std::to_string(threads) + "/" + g_version;
  if (out && cap) {
    const size_t n = std::min(cap - 1, key.size());
    std::memcpy(out, key.data(), n);
    out[n] = 0;
  }
  return key.size();
}

size_t og_tuning_choose(const double *rate, const uint64_t *bytes, size_t n, uint64_t budget,
                        double tolerance) {
  size_t best = (size_t)-1;
  for (size_t i = 0; i < n; ++i)
    if (rate[i] > 0 && (!budget || bytes[i] <= budget) && (best == (size_t)-1 || rate[i] > rate[best]))
      best = i;
  if (best == (size_t)-1) return best;
  const double floor = rate[best] * (1.0 - tolerance);
  size_t pick = best;
  for (size_t i = 0; i < n; ++i)
    if (rate[i] > 0 && rate[i] >= floor && (!budget || bytes[i] <= budget) &&
        (bytes[i] < bytes[pick] || (bytes[i] == bytes[pick] && rate[i] > rate[pick])))
      pick = i;
  return pick;
}

int og_tuning_path(char *out, size_t cap) {
  std::string p = cache_path();
  if (p.empty() || !out || cap <= p.size()) return -1;
  std::memcpy(out, p.c_str(), p.size() + 1);
  return 0;
}

int og_tuning_get(const char *key, og_tuning_shape *out) {
  if (!key || !out) return 0;
  std::map<std::string, og_tuning_shape> m;
  if (!read_file(m)) return 0;
  auto it = m.find(key);
  if (it == m.end()) return 0;
  *out = it->second;
  return 1;
}

int og_tuning_put(const char *key, const og_tuning_shape *shape) {
  if (!key || !shape) return -1;
  std::string path = cache_path();
  if (path.empty()) return -1;
  std::map<std::string, og_tuning_shape> m;
  if (!read_file(m)) m.clear();  // a foreign format is replaced, never misread
  og_tuning_shape s = *shape;
  if (!s.at) s.at = (int64_t)std::time(nullptr);
  m[key] = s;
  std::ostringstream o;
  o << "format: 1\nentries:\n";
  for (auto &kv : m) {
    o << "  '" << kv.first << "':\n";
    o << "    first: " << kv.second.first << "\n";
    o << "    second: " << kv.second.second << "\n";
    char rate[64];
    std::snprintf(rate, sizeof rate, "%.6g", kv.second.rate);
    o << "    rate: " << rate << "\n";
    o << "    at: " << kv.second.at << "\n";
    if (kv.second.rounds) o << "    rounds: " << kv.second.rounds << "\n";
    if (kv.second.radix) o << "    radix: " << kv.second.radix << "\n";
    if (kv.second.slots) o << "    slots: " << kv.second.slots << "\n";
  }
  mkdirs(path);
  // written beside and renamed, so a reader never sees half a file
  std::string tmp = path + ".new";
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return -1;
    f << o.str();
    if (!f) return -1;
  }
  return std::rename(tmp.c_str(), path.c_str()) == 0 ? 0 : -1;
}

}  // extern "C"
