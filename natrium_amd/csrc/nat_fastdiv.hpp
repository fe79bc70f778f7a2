// Unsigned 32-bit division by a divisor fixed at launch time, as a multiply-high and
// shifts (T. Granlund, P. Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4): for 1 <= d < 2^31, l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, and for every 32-bit n
//   t = mulhi(n, m),  n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
// A division by a run-time value otherwise expands to ~30 VALU instructions; the
// kernels use it for cell indices -> cell coordinates once per (DoF, direction).
// Plain C++ on the host (no HIP headers needed), so the host builds the constants and
// CPU tests check the identity.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define NAT_HD __host__ __device__
#else
#define NAT_HD
#endif

namespace nat {

struct FastDivU {
  uint32_t m = 1;
  uint32_t s1 = 0, s2 = 0;
  uint32_t d = 1;

  FastDivU() = default;
  explicit FastDivU(uint32_t div) { set(div); }

  void set(uint32_t div) {
    d = div < 1 ? 1 : div;
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) l++;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    s1 = l < 1 ? l : 1;
    s2 = l > 1 ? l - 1 : 0;
  }

  NAT_HD uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t t = __umulhi(n, m);
#else
    const uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
    return (t + ((n - t) >> s1)) >> s2;
  }
};

}  // namespace nat
