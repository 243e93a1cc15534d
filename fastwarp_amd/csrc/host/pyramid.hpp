// Multi-resolution pyramid of antsRegistration (crates/fastwarp-reg/src/pyramid.rs): ANTs shrink
// factors per axis, ITK ShrinkImageFilter virtual grids, and ITK's recursive Gaussian (Deriche,
// fourth order, constant edge extension) of the full-resolution images, in f64 on the host.
#pragma once

#include "grid.hpp"

namespace fw {

enum class SigmaUnits { Voxels, Millimetres };

// pyramid.rs:63-92 `shrink_factors_per_dimension`
inline std::array<uint32_t, 3> shrink_factors_per_dimension(uint32_t factor, const Vec3& spacing) {
  factor = std::max<uint32_t>(factor, 1);
  int min_index = 0;
  double min_spacing = spacing[0];
  for (int n = 1; n < 3; ++n)
    if (min_spacing > spacing[n]) {
      min_spacing = spacing[n];
      min_index = n;
    }
  const double target = spacing[min_index] * double(factor);
  std::array<uint32_t, 3> out{0, 0, 0};
  out[min_index] = factor;
  for (int n = 0; n < 3; ++n) {
    if (out[n] != 0) continue;
    double best_diff = std::fabs(spacing[n] * double(factor) - target);
    uint32_t best = factor;
    for (uint32_t f = factor - 1; f >= 1; --f) {
      const double diff = std::fabs(spacing[n] * double(f) - target);
      if (diff <= best_diff) {
        best_diff = diff;
        best = f;
      }
      if (f == 1) break;
    }
    out[n] = best;
  }
  return out;
}

// pyramid.rs:100-129 `shrink_grid`: floor sizes (at least 1), spacing times factor, origin moved
// so the physical centres coincide.
inline Grid shrink_grid(const Grid& g, const std::array<uint32_t, 3>& factors) {
  std::array<size_t, 3> size{};
  Vec3 spacing{};
  for (int a = 0; a < 3; ++a) {
    const uint32_t f = std::max<uint32_t>(factors[a], 1);
    spacing[a] = g.spacing[a] * double(f);
    size[a] = std::max<size_t>(size_t(std::floor(double(g.size[a]) / double(f))), 1);
  }
  const Vec3 in_c{(double(g.size[0]) - 1) / 2, (double(g.size[1]) - 1) / 2, (double(g.size[2]) - 1) / 2};
  const Vec3 out_c{(double(size[0]) - 1) / 2, (double(size[1]) - 1) / 2, (double(size[2]) - 1) / 2};
  const Vec3 pi = g.index_to_point(in_c);
  const Grid provisional(size, spacing, g.origin, g.direction);
  const Vec3 po = provisional.index_to_point(out_c);
  const Vec3 origin{g.origin[0] + (pi[0] - po[0]), g.origin[1] + (pi[1] - po[1]), g.origin[2] + (pi[2] - po[2])};
  return Grid(size, spacing, origin, g.direction);
}

inline Grid level_virtual_grid(const Grid& fixed, const std::array<uint32_t, 3>& factors) { return shrink_grid(fixed, factors); }

// pyramid.rs:184-305: Deriche's fourth-order recursive approximation of a zero-order Gaussian.
struct Deriche {
  double n[4], d[4], m[4], bn[4], bm[4];

  static Deriche zero_order(double sigma, double spacing) {
    if (spacing < 1e-8) invalid("smoothing", "spacing is too small");
    const double sd = sigma / spacing;
    const double a1 = 1.3530, b1 = 1.8151, w1 = 0.6681, l1 = -1.3932;
    const double a2 = -0.3531, b2 = 0.0902, w2 = 2.0787, l2 = -1.3732;
    const double sin1 = std::sin(w1 / sd), sin2 = std::sin(w2 / sd);
    const double cos1 = std::cos(w1 / sd), cos2 = std::cos(w2 / sd);
    const double e1 = std::exp(l1 / sd), e2 = std::exp(l2 / sd);
    Deriche k{};
    const double d4 = e1 * e1 * e2 * e2;
    double d3 = -2.0 * cos1 * e1 * e2 * e2;
    d3 += -2.0 * cos2 * e2 * e1 * e1;
    double d2 = 4.0 * cos2 * cos1 * e1 * e2;
    d2 += e1 * e1 + e2 * e2;
    const double d1 = -2.0 * (e2 * cos2 + e1 * cos1);
    const double sdd = 1.0 + d1 + d2 + d3 + d4;
    double n0 = a1 + a2;
    double n1 = e2 * (b2 * sin2 - (a2 + 2.0 * a1) * cos2);
    n1 += e1 * (b1 * sin1 - (a1 + 2.0 * a2) * cos1);
    double n2 = (a1 + a2) * cos2 * cos1;
    n2 -= b1 * cos2 * sin1 + b2 * cos1 * sin2;
    n2 *= 2.0 * e1 * e2;
    n2 += a2 * e1 * e1 + a1 * e2 * e2;
    double n3 = e2 * e1 * e1 * (b2 * sin2 - a2 * cos2);
    n3 += e1 * e2 * e2 * (b1 * sin1 - a1 * cos1);
    const double sn0 = n0 + n1 + n2 + n3;
    const double alpha0 = 2.0 * sn0 / sdd - n0;
    n0 /= alpha0;
    n1 /= alpha0;
    n2 /= alpha0;
    n3 /= alpha0;
    const double m1 = n1 - d1 * n0, m2 = n2 - d2 * n0, m3 = n3 - d3 * n0, m4 = -d4 * n0;
    const double sn = n0 + n1 + n2 + n3;
    const double sm = m1 + m2 + m3 + m4;
    const double dd[4] = {d1, d2, d3, d4};
    const double nn[4] = {n0, n1, n2, n3};
    const double mm[4] = {m1, m2, m3, m4};
    for (int q = 0; q < 4; ++q) {
      k.n[q] = nn[q];
      k.d[q] = dd[q];
      k.m[q] = mm[q];
      k.bn[q] = dd[q] * sn / sdd;
      k.bm[q] = dd[q] * sm / sdd;
    }
    return k;
  }

  // pyramid.rs:246-304: causal and anti-causal passes with constant extension; line length >= 4.
  void filter_line(const double* data, size_t ln, double* causal, double* anti, double* out) const {
    const double v1 = data[0];
    causal[0] = v1 * n[0] + v1 * n[1] + v1 * n[2] + v1 * n[3];
    causal[1] = data[1] * n[0] + v1 * n[1] + v1 * n[2] + v1 * n[3];
    causal[2] = data[2] * n[0] + data[1] * n[1] + v1 * n[2] + v1 * n[3];
    causal[3] = data[3] * n[0] + data[2] * n[1] + data[1] * n[2] + v1 * n[3];
    causal[0] -= v1 * bn[0] + v1 * bn[1] + v1 * bn[2] + v1 * bn[3];
    causal[1] -= causal[0] * d[0] + v1 * bn[1] + v1 * bn[2] + v1 * bn[3];
    causal[2] -= causal[1] * d[0] + causal[0] * d[1] + v1 * bn[2] + v1 * bn[3];
    causal[3] -= causal[2] * d[0] + causal[1] * d[1] + causal[0] * d[2] + v1 * bn[3];
    for (size_t i = 4; i < ln; ++i) {
      causal[i] = data[i] * n[0] + data[i - 1] * n[1] + data[i - 2] * n[2] + data[i - 3] * n[3];
      causal[i] -= causal[i - 1] * d[0] + causal[i - 2] * d[1] + causal[i - 3] * d[2] + causal[i - 4] * d[3];
    }
    const double v2 = data[ln - 1];
    anti[ln - 1] = v2 * m[0] + v2 * m[1] + v2 * m[2] + v2 * m[3];
    anti[ln - 2] = data[ln - 1] * m[0] + v2 * m[1] + v2 * m[2] + v2 * m[3];
    anti[ln - 3] = data[ln - 2] * m[0] + data[ln - 1] * m[1] + v2 * m[2] + v2 * m[3];
    anti[ln - 4] = data[ln - 3] * m[0] + data[ln - 2] * m[1] + data[ln - 1] * m[2] + v2 * m[3];
    anti[ln - 1] -= v2 * bm[0] + v2 * bm[1] + v2 * bm[2] + v2 * bm[3];
    anti[ln - 2] -= anti[ln - 1] * d[0] + v2 * bm[1] + v2 * bm[2] + v2 * bm[3];
    anti[ln - 3] -= anti[ln - 2] * d[0] + anti[ln - 1] * d[1] + v2 * bm[2] + v2 * bm[3];
    anti[ln - 4] -= anti[ln - 3] * d[0] + anti[ln - 2] * d[1] + anti[ln - 1] * d[2] + v2 * bm[3];
    for (size_t i = ln - 4; i >= 1; --i) {
      anti[i - 1] = data[i] * m[0] + data[i + 1] * m[1] + data[i + 2] * m[2] + data[i + 3] * m[3];
      anti[i - 1] -= anti[i] * d[0] + anti[i + 1] * d[1] + anti[i + 2] * d[2] + anti[i + 3] * d[3];
    }
    for (size_t i = 0; i < ln; ++i) out[i] = causal[i] + anti[i];
  }
};

// Filters every line of `data` along `axis` in place, lines in parallel (pyramid.rs:366-419).
inline void deriche_axis(HostVec<double>& data, const std::array<size_t, 3>& size, int axis, const Deriche& k) {
  const size_t nx = size[0], ny = size[1], nz = size[2];
  const size_t ln = size[axis];
  const size_t stride = axis == 0 ? 1 : axis == 1 ? nx : nx * ny;
  const size_t lines = data.size() / ln;
  parallel_chunks(lines, 64, [&](size_t, size_t b, size_t e) {
    std::vector<double> line(ln), causal(ln), anti(ln), out(ln);
    for (size_t l = b; l < e; ++l) {
      size_t base;
      if (axis == 0) base = l * nx;
      else if (axis == 1) base = (l / nx) * nx * ny + (l % nx);
      else base = l;  // l < nx*ny
      for (size_t t = 0; t < ln; ++t) line[t] = data[base + t * stride];
      k.filter_line(line.data(), ln, causal.data(), anti.data(), out.data());
      for (size_t t = 0; t < ln; ++t) data[base + t * stride] = out[t];
    }
  });
  (void)nz;
}

// pyramid.rs:429-437: axes z, x, y with per-axis sigma in mm; every axis needs 4 voxels.
inline Image smooth_recursive_gaussian(const Image& vol, const Vec3& sigma_mm) {
  HostVec<double> out = vol.data;
  for (int axis : {2, 0, 1}) {
    if (vol.grid.size[axis] < 4) invalid("smoothing", "every axis needs at least 4 voxels");
    if (!(sigma_mm[axis] > 0.0)) invalid("smoothing", "sigma must be greater than zero");
    deriche_axis(out, vol.grid.size, axis, Deriche::zero_order(sigma_mm[axis], vol.grid.spacing[axis]));
  }
  return Image(vol.grid, std::move(out));
}

// pyramid.rs:441-465 `sigma_mm_per_axis`, `smooth_for_level`: a copy for sigma <= 0.
inline Image smooth_for_level(const Image& vol, double sigma, SigmaUnits units) {
  HostTimer host_timer("deriche");
  if (sigma <= 0.0) return vol;
  const Vec3& s = vol.grid.spacing;
  const Vec3 mm = units == SigmaUnits::Voxels ? Vec3{sigma * s[0], sigma * s[1], sigma * s[2]} : Vec3{sigma, sigma, sigma};
  return smooth_recursive_gaussian(vol, mm);
}

}  // namespace fw
