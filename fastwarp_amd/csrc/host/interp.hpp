// Host-side linear interpolation and central differences with ITK's buffer rules
// (crates/fastwarp-reg/src/interp.rs). The product uses them for the work the reference also
// keeps on the host: the fixed bin of each linear-stage sample, centres of mass and resampling.
#pragma once

#include "grid.hpp"

namespace fw {

// interp.rs:43-87: ITK LinearInterpolateImageFunction::EvaluateAtContinuousIndex (base index
// clamped to the buffer; no interpolation across an axis whose distance is not positive or whose
// upper neighbour is past the last voxel).
inline double linear_at_index(const Image& vol, const Vec3& c) {
  const auto& size = vol.grid.size;
  size_t base[3];
  double w[3];
  for (int a = 0; a < 3; ++a) {
    const double f = std::floor(c[a]);
    size_t b = f < 0.0 ? 0 : static_cast<size_t>(f);
    b = std::min(b, size[a] - 1);
    const double d = c[a] - static_cast<double>(b);
    base[a] = b;
    w[a] = (d <= 0.0 || b + 1 > size[a] - 1) ? 0.0 : d;
  }
  auto at = [&](size_t dx, size_t dy, size_t dz) {
    return vol.data[vol.grid.linear_index(base[0] + dx, base[1] + dy, base[2] + dz)];
  };
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  auto xr = [&](size_t dy, size_t dz) {
    const double v0 = at(0, dy, dz);
    return w[0] > 0.0 ? lerp(v0, at(1, dy, dz), w[0]) : v0;
  };
  auto yr = [&](size_t dz) {
    const double v0 = xr(0, dz);
    return w[1] > 0.0 ? lerp(v0, xr(1, dz), w[1]) : v0;
  };
  const double v0 = yr(0);
  return w[2] > 0.0 ? lerp(v0, yr(1), w[2]) : v0;
}

// interp.rs:90-95; false outside the buffer.
inline bool linear_at_point(const Image& vol, const Vec3& p, double& out) {
  const Vec3 c = vol.grid.point_to_index(p);
  if (!vol.grid.contains_continuous_index(c)) return false;
  out = linear_at_index(vol, c);
  return true;
}

// mask.rs:99-111: nearest voxel rounding half up; false outside the grid.
inline bool nearest_voxel(const Grid& g, const Vec3& p, size_t& out) {
  const Vec3 c = g.point_to_index(p);
  size_t idx[3];
  for (int a = 0; a < 3; ++a) {
    const double r = std::floor(c[a] + 0.5);
    if (!(r >= 0.0 && r < static_cast<double>(g.size[a]))) return false;
    idx[a] = static_cast<size_t>(r);
  }
  out = g.linear_index(idx[0], idx[1], idx[2]);
  return true;
}

}  // namespace fw
