// Division of 32-bit indices by a runtime divisor with one multiply-high, for kernels that turn a
// linear voxel index into (i, j, k): a 64-bit division costs a long instruction sequence on the GPU.
// Granlund and Montgomery, "Division by invariant integers using multiplication" (PLDI 1994), the
// round-up variant: with s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1,
// floor(n / d) = (mulhi(n, m) + n) >> s for every n < 2^31 (the sum then stays below 2^32).
#pragma once

#include <cstdint>

namespace fw {

struct FastDiv {
  uint32_t d, m, s;
};

// d >= 1.
constexpr FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
  return FastDiv{d, uint32_t(m), s};
}

// floor(n / f.d) for n < 2^31 (constexpr functions run on the host and in kernels alike).
constexpr uint32_t fast_div(uint32_t n, FastDiv f) {
  const uint32_t t = uint32_t((uint64_t(n) * f.m) >> 32);
  return (t + n) >> f.s;
}

}  // namespace fw
