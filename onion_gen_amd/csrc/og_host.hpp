// Host-side curve constants and tables the kernels read.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/onion_gen_hip.h"
#include "ge.hpp"
#include "og_filters.hpp"
#include "og_kernels.hpp"

namespace og {

struct HostCurve {
  Fe d, two_d;
  Point base;  // B (src/curve.rs:28-31)

  HostCurve();
  Cached eight_cached() const;  // 8G
  // Q_m = m * 8G for m = 1..half as (x, y, d*x*y), canonical, 24 words each.
  std::vector<uint32_t> offset_table(uint32_t half) const;
  // The kernels' table in the 10-limb form: per entry xq, yq, dq and the
  // derived terms of offset_terms, kTableWords words.
  std::vector<uint32_t> offset_table10(uint32_t half) const;
  // Set-up arguments for chains [first, first + count) of a start point given
  // as 128 bytes xyzt, in a state buffer of SoA stride `threads`; chain c's
  // first window starts at candidate k0 + c(2*half+1).
  SetupArgs setup_args(const uint8_t start_xyzt[128], uint32_t *state, uint32_t threads,
                       uint32_t first, uint32_t count, uint32_t half, uint64_t k0 = 0) const;
  // m * 8G with Z scaled to one.
  Cached1 multiple_of_eight_cached1(uint64_t m) const;
  // scalar * B, 256-bit scalar, double-and-add from the top bit
  // (src/curve.rs:188-198).
  Point scalar_base_mult(const uint8_t scalar[32]) const;
  void pack(const Point &p, uint8_t out[32]) const;  // src/key.rs:55-62
};

const FilterSet *filterset_of(const og_filterset *s);

// Writes the 10-limb form of an 8-word value.
void put_fe10(uint32_t out[10], const Fe &x);

}  // namespace og
