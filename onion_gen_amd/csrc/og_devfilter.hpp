// Host side: a FilterSet compiled into the tables a search kernel reads, and
// their device copies.  Shared by the single-search engine and the slot
// engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "og_filters.hpp"
#include "og_kernels.hpp"

namespace og {

struct DevFilterHost {
  uint32_t mode = 0;
  uint32_t want[8] = {0}, mask[8] = {0};
  uint32_t bits = 0;
  std::vector<uint32_t> bitmap;   // 32-bit halves of the 64-bit words
  std::vector<uint32_t> buckets;  // [offset, count, mask_lo, mask_hi]
  std::vector<uint64_t> keys;
  std::vector<uint32_t> pats;     // kPatWords per pattern
  // device copies
  uint32_t *d_bitmap = nullptr, *d_buckets = nullptr, *d_pats = nullptr;
  uint64_t *d_keys = nullptr;

  size_t device_bytes() const {
    return bitmap.size() * 4 + buckets.size() * 4 + keys.size() * 8 + pats.size() * 4;
  }

  // Empty string on success.
  std::string compile(const FilterSet &fs) {
    if (fs.fast.size() == 1) {
      const Filter &f = fs.filters[fs.fast[0]];
      mode |= kModePrefix;
      for (int i = 0; i < f.nbytes; ++i) {
        const uint8_t m = i == f.nbytes - 1 ? f.mask : 0xff;
        want[i / 4] |= (uint32_t)f.bytes[i] << (8 * (i % 4));
        mask[i / 4] |= (uint32_t)m << (8 * (i % 4));
      }
    } else if (!fs.fast.empty()) {
      mode |= kModeBitmap;
      // The reference's index when it built one; otherwise a 24-bit map at
      // any occupancy, which is still a valid "maybe" for the device.
      const std::vector<uint64_t> w = fs.index_bits ? fs.words : bitmap_for(fs, 24);
      bits = fs.index_bits ? fs.index_bits : 24;
      set_bitmap(w);
      if (fs.buckets.size() > (size_t)kMaxBuckets)
        return "more distinct prefix lengths than the device table holds";
      for (const Bucket &b : fs.buckets) {
        buckets.push_back((uint32_t)keys.size());
        buckets.push_back((uint32_t)b.keys.size());
        buckets.push_back((uint32_t)b.mask);
        buckets.push_back((uint32_t)(b.mask >> 32));
        keys.insert(keys.end(), b.keys.begin(), b.keys.end());
      }
    }
    if (!fs.general.empty()) {
      if (fs.general.size() > (size_t)kMaxPatterns)
        return "at most 8 general (suffix, contains, class) filters per search";
      mode |= kModePattern;
      for (uint32_t i : fs.general) {
        DevPattern p = dev_pattern(fs.filters[i]);
        pats.push_back(p.len);
        pats.push_back(p.first);
        pats.push_back(p.last);
        pats.push_back(p.key_last);
        for (int s = 0; s < kAddressLen; ++s) pats.push_back(p.sets[s]);
      }
    }
    if (!mode) return "the filter set is empty";
    return "";
  }

  void set_bitmap(const std::vector<uint64_t> &w) {
    bitmap.clear();
    for (uint64_t v : w) {
      bitmap.push_back((uint32_t)v);
      bitmap.push_back((uint32_t)(v >> 32));
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
    if (e == hipSuccess) e = upload(&d_buckets, buckets, s);
    if (e == hipSuccess) e = upload(&d_keys, keys, s);
    if (e == hipSuccess) e = upload(&d_pats, pats, s);
    return e;
  }

  void free_device() {
    (void)hipFree(d_bitmap);
    (void)hipFree(d_buckets);
    (void)hipFree(d_keys);
    (void)hipFree(d_pats);
    d_bitmap = d_buckets = d_pats = nullptr;
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
    a.nbuckets = (uint32_t)(buckets.size() / 4);
    a.buckets = d_buckets;
    a.keys = d_keys;
    a.npat = (uint32_t)(pats.size() / kPatWords);
    a.pats = d_pats;
    return a;
  }
};

}  // namespace og
