// ITK's Mersenne Twister and the REGULAR metric sampling of ImageRegistrationMethodv4
// (crates/fastwarp-reg/src/sampling.rs:44-234).
#pragma once

#include <cstdint>

#include "grid.hpp"

namespace fw {

// The first n outputs of Mt19937(seed) before tempering, as one array: with x[0..623] the seeded
// state, x[k] = x[k - 227] ^ twist(x[k - 624], x[k - 623]) for k >= 624 are the state words that the
// generator's reloads produce in order, and its i-th output is mt19937_temper(x[624 + i]). The
// recurrence reaches 227 words back at the least, so the loop runs in vector lanes; tempering is
// left to the caller's parallel loop.
inline HostVec<uint32_t> mt19937_untempered(uint32_t seed, size_t n) {
  constexpr size_t kN = 624, kGap = 624 - 397;
  HostVec<uint32_t> x(kN + n);
  x[0] = seed;
  for (uint32_t i = 1; i < kN; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + i;
  uint32_t* p = x.data();
  for (size_t k = kN; k < kN + n; ++k) {
    const uint32_t mix = (p[k - kN] & 0x80000000u) | (p[k - kN + 1] & 0x7fffffffu);
    p[k] = p[k - kGap] ^ (mix >> 1) ^ ((p[k - kN + 1] & 1u) ? 0x9908b0dfu : 0u);
  }
  return x;
}

inline uint32_t mt19937_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  return y ^ (y >> 18);
}

// sampling.rs:44-121: MT19937 seeded with a 32-bit integer, ITK's reload/twist/tempering.
class Mt19937 {
 public:
  explicit Mt19937(uint32_t seed) {
    s_[0] = seed;
    for (uint32_t i = 1; i < kN; ++i) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + i;
    reload();
  }
  uint32_t integer() {
    if (left_ == 0) reload();
    --left_;
    return mt19937_temper(s_[next_++]);
  }

 private:
  static constexpr uint32_t kN = 624, kM = 397;
  static uint32_t twist(uint32_t m, uint32_t s0, uint32_t s1) {
    const uint32_t mix = (s0 & 0x80000000u) | (s1 & 0x7fffffffu);
    return m ^ (mix >> 1) ^ ((s1 & 1u) ? 0x9908b0dfu : 0u);
  }
  void reload() {
    for (uint32_t p = 0; p < kN - kM; ++p) s_[p] = twist(s_[p + kM], s_[p], s_[p + 1]);
    for (uint32_t p = kN - kM; p < kN - 1; ++p) s_[p] = twist(s_[p + kM - kN], s_[p], s_[p + 1]);
    s_[kN - 1] = twist(s_[kM - 1], s_[kN - 1], s_[0]);
    left_ = kN;
    next_ = 0;
  }
  uint32_t s_[kN];
  uint32_t next_ = 0, left_ = 0;
};

// sampling.rs:173-181: ITK's Box-Muller from one open-range and one open-upper draw, cosine only.
inline double normal_from(uint32_t open, uint32_t open_upper, double mean, double variance) {
  const double u1 = (double(open) + 0.5) * (1.0 / 4294967296.0);
  const double u2 = double(open_upper) * (1.0 / 4294967296.0);
  const double r = std::sqrt(-2.0 * std::log(1.0 - u1) * variance);
  return mean + r * std::cos(2.0 * 3.14159265358979323846 * u2);
}

// sampling.rs:143-164: the per-stage seed sequence, incremented once per generator.
struct SeedSequence {
  uint32_t current;
  explicit SeedSequence(uint32_t seed) : current(seed) {}
  uint32_t peek() const { return current; }
  uint32_t next_seed() { return current++; }
};

// sampling.rs:196-234: voxel 0 and every ceil(1 / percentage)-th voxel in raster order, each
// jittered by N(0, 1) * spacing / 3 per axis (x first); `keep` filters after the jitter.
inline HostVec<Vec3> metric_sample_points(const Grid& g, float percentage, SeedSequence& seeds) {
  HostTimer host_timer("sample_points");
  const uint32_t seed = seeds.next_seed();
  const size_t stride = size_t(std::max(1.0, std::ceil(1.0 / double(percentage))));
  const size_t n = g.num_voxels();
  const size_t picks = (n + stride - 1) / stride;
  const HostVec<uint32_t> raw = mt19937_untempered(seed, 6 * picks);
  const uint32_t* draws = raw.data() + 624;
  HostVec<Vec3> pts(picks);
  parallel_for(picks, [&](size_t b, size_t e) {
    for (size_t q = b; q < e; ++q) {
      const size_t flat = q * stride;
      auto [i, j, k] = g.ijk(flat);
      Vec3 p = g.index_to_point({double(i), double(j), double(k)});
      for (int d = 0; d < 3; ++d)
        p[d] += normal_from(mt19937_temper(draws[6 * q + 2 * d]), mt19937_temper(draws[6 * q + 2 * d + 1]), 0.0, 1.0) *
                (g.spacing[d] / 3.0);
      pts[q] = p;
    }
  });
  return pts;
}

template <typename Keep>
HostVec<Vec3> metric_sample_points(const Grid& g, float percentage, SeedSequence& seeds, Keep keep) {
  const HostVec<Vec3> all = metric_sample_points(g, percentage, seeds);
  HostVec<Vec3> out;
  out.reserve(all.size());
  for (const Vec3& p : all)
    if (keep(p)) out.push_back(p);
  return out;
}

}  // namespace fw
