// Writing a verified key pair in the layout tor reads (src/output.rs:15-88):
// a directory named after the hostname, mode 0700, holding
//   hs_ed25519_secret_key  "== ed25519v1-secret: type0 ==" + 3 zeros + 64-byte expanded secret (0600)
//   hs_ed25519_public_key  "== ed25519v1-public: type0 ==" + 3 zeros + 32-byte key
//   hostname               "<address>.onion\n"
// An existing directory of that name is left untouched.
#include <errno.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <string>

#include "../../include/onion_gen_hip.h"

namespace {

void put_path(char *out, size_t cap, const std::string &s) {
  if (out && cap) {
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(out, s.data(), n);
    out[n] = 0;
  }
}

bool write_file(const std::string &path, const uint8_t *data, size_t n, mode_t mode) {
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_EXCL, mode);
  if (fd < 0) return false;
  size_t done = 0;
  while (done < n) {
    const ssize_t w = ::write(fd, data + done, n - done);
    if (w <= 0) {
      ::close(fd);
      return false;
    }
    done += (size_t)w;
  }
  if (::fchmod(fd, mode) != 0) {  // the process umask must not widen it
    ::close(fd);
    return false;
  }
  return ::close(fd) == 0;
}

bool make_dirs(const std::string &dir) {
  std::string cur;
  for (size_t i = 0; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      cur = dir.substr(0, i);
      if (!cur.empty() && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
    }
  }
  return true;
}

}  // namespace

extern "C" int og_write_key(const char *dir, const uint8_t public_key[32],
                            const uint8_t expanded_secret[64], char *path_out, size_t cap) {
  static const char kSecret[32] = "== ed25519v1-secret: type0 ==\0\0";
  static const char kPublic[32] = "== ed25519v1-public: type0 ==\0\0";
  if (!dir || !public_key || !expanded_secret) {
    put_path(path_out, cap, "null argument");
    return -1;
  }
  char addr[57];
  og_key_address(public_key, addr);
  const std::string hostname = std::string(addr) + ".onion";
  const std::string target = std::string(dir) + "/" + hostname;
  if (!make_dirs(dir)) {
    put_path(path_out, cap, std::string("cannot create ") + dir + ": " + std::strerror(errno));
    return -1;
  }
  if (::mkdir(target.c_str(), 0700) != 0) {
    if (errno == EEXIST) {
      put_path(path_out, cap, target);
      return 1;
    }
    put_path(path_out, cap, "cannot create " + target + ": " + std::strerror(errno));
    return -1;
  }
  if (::chmod(target.c_str(), 0700) != 0) {
    put_path(path_out, cap, "cannot restrict " + target + ": " + std::strerror(errno));
    return -1;
  }
  uint8_t secret[96], pub[64];
  std::memcpy(secret, kSecret, 32);
  std::memcpy(secret + 32, expanded_secret, 64);
  std::memcpy(pub, kPublic, 32);
  std::memcpy(pub + 32, public_key, 32);
  const std::string host_line = hostname + "\n";
  if (!write_file(target + "/hs_ed25519_secret_key", secret, sizeof secret, 0600) ||
      !write_file(target + "/hs_ed25519_public_key", pub, sizeof pub, 0644) ||
      !write_file(target + "/hostname", (const uint8_t *)host_line.data(), host_line.size(),
                  0644)) {
    put_path(path_out, cap, "cannot write the key files under " + target + ": " +
                                std::strerror(errno));
    return -1;
  }
  put_path(path_out, cap, target);
  return 0;
}
