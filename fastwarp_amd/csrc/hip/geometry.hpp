// Host-side geometry for the launches: grids to dims, f64 maps narrowed to f32, probe offsets.
#pragma once

#include <cmath>

#include "../host/grid.hpp"
#include "kernels.hpp"

namespace fw::hip {

inline Dims dims_of(const Grid& g) { return {int(g.size[0]), int(g.size[1]), int(g.size[2])}; }

inline Mat3 scaled_direction(const Grid& g) {
  Mat3 m = g.direction;
  for (auto& row : m)
    for (int c = 0; c < 3; ++c) row[c] *= g.spacing[c];
  return m;
}

inline void put(float* dst, const Mat3& m) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dst[r * 3 + c] = float(m[r][c]);
}

inline void put(float* dst, const Vec3& v) {
  for (int r = 0; r < 3; ++r) dst[r] = float(v[r]);
}

inline float3 inv_spacing(const Grid& g) {
  return make_float3(float(1.0 / g.spacing[0]), float(1.0 / g.spacing[1]), float(1.0 / g.spacing[2]));
}

// The central-difference probe offsets of a grid in index units: +-half a spacing along each
// physical axis d maps to index_from_offset * (h_d e_d) (interp.rs:102-121).
inline Probes probes_of(const Grid& g) {
  Probes p{};
  bool diagonal = true;
  for (int d = 0; d < 3; ++d) {
    const double h = 0.5 * g.spacing[d];
    int nonzero = 0;
    for (int r = 0; r < 3; ++r) {
      double v = g.index_from_offset[r][d] * h;
      if (std::fabs(v) < 1e-12) v = 0.0;
      if (v != 0.0) {
        ++nonzero;
        if (std::fabs(std::fabs(v) - 0.5) < 1e-9) v = v > 0 ? 0.5 : -0.5;
        else diagonal = false;
      }
      p.o[d][r] = float(v);
    }
    if (nonzero != 1) diagonal = false;
    p.inv_spacing[d] = float(1.0 / g.spacing[d]);
  }
  p.diagonal = diagonal ? 1 : 0;
  return p;
}


// The split scale of the CC sums for an image bounded by `bound`: 2^19 / 2^e with 2^e >= bound,
// so that |v| * scale <= 2^19 (2^20 if the bound is exceeded up to twice).
inline float cc_split_scale(float bound) {
  if (!(bound > 0.0f) || !std::isfinite(bound)) return float(1 << 19);
  int e = 0;
  std::frexp(double(bound), &e);  // bound < 2^e
  return float(std::ldexp(1.0, 19 - e));
}

}  // namespace fw::hip
