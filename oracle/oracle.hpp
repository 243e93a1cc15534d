// ORACLE — test infrastructure only. This is a CPU f64 restatement of the reference's CPU engine
// (`engine="cpu"`: crates/fastwarp-reg/src/cpu/mod.rs `CpuBackend`, metric/cc.rs, metric/mi.rs,
// field_ops.rs, interp.rs). It is the parity checker of the HIP path and the timed CPU baseline;
// only the tests, the smoke check and the CPU-baseline leg of the benchmark may load it. The
// product never calls into it.
//
// It shares with the product only the plain data types (grid, volume, field, transforms) and the
// generic host drivers (SyN loop, linear stage, optimiser), as the reference's CPU and GPU
// engines share `syn_fields` and `run_linear_stage` over the `Backend`/`LinearEngine` traits.
// Every numeric operation below is written here, independently of the HIP kernels.
//
// Parity pins (tests/test_oracle.py): the ITK 5.4.5 Mattes fixtures (dense MI value, valid count
// and derivative, masked and oblique), CC(I,I) = -1 with zero derivative, CC against a brute-force
// window on 9^3, the CC coefficient against finite differences, affine-intensity invariance, the
// Gaussian taps against the Bessel series, delta and constant smoothing, closed-form compose,
// resample and inversion cases, B-spline partition of unity, and MI parameter derivatives against
// finite differences of the value.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <optional>
#include <utility>
#include <vector>

#include "../fastwarp_amd/csrc/host/grid.hpp"
#include "../fastwarp_amd/csrc/host/pyramid.hpp"
#include "../fastwarp_amd/csrc/host/transform.hpp"

namespace fw::oracle {

// ---------------------------------------------------------------- interpolation (interp.rs) --

// interp.rs:43-87 (ITK EvaluateOptimized in 3D).
inline double lerp_index(const Image& vol, const Vec3& c) {
  const auto& n = vol.grid.size;
  size_t base[3];
  double w[3];
  for (int a = 0; a < 3; ++a) {
    const double f = std::floor(c[a]);
    size_t b = f < 0.0 ? 0 : size_t(f);
    if (b > n[a] - 1) b = n[a] - 1;
    const double d = c[a] - double(b);
    base[a] = b;
    w[a] = (d <= 0.0 || b + 1 > n[a] - 1) ? 0.0 : d;
  }
  const auto& g = vol.grid;
  auto at = [&](size_t dx, size_t dy, size_t dz) { return vol.data[g.linear_index(base[0] + dx, base[1] + dy, base[2] + dz)]; };
  auto x = [&](size_t dy, size_t dz) {
    double v0 = at(0, dy, dz);
    return w[0] > 0.0 ? v0 + (at(1, dy, dz) - v0) * w[0] : v0;
  };
  auto y = [&](size_t dz) {
    double v0 = x(0, dz);
    return w[1] > 0.0 ? v0 + (x(1, dz) - v0) * w[1] : v0;
  };
  double v0 = y(0);
  return w[2] > 0.0 ? v0 + (y(1) - v0) * w[2] : v0;
}

// interp.rs:90-95
inline std::optional<double> lerp_point(const Image& vol, const Vec3& p) {
  const Vec3 c = vol.grid.point_to_index(p);
  if (!vol.grid.contains_continuous_index(c)) return std::nullopt;
  return lerp_index(vol, c);
}

// interp.rs:102-121: ITK CentralDifferenceImageFunction at a point, along the physical axes.
inline Vec3 central_difference(const Image& vol, const Vec3& p) {
  Vec3 out{0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    const double off = 0.5 * vol.grid.spacing[d];
    Vec3 p1 = p, p2 = p;
    p1[d] = p[d] - off;
    p2[d] = p[d] + off;
    auto v1 = lerp_point(vol, p1);
    auto v2 = lerp_point(vol, p2);
    if (!v1 || !v2) continue;
    const double delta = p2[d] - p1[d];
    if (delta > 10.0 * kEps) out[d] = (*v2 - *v1) / delta;
  }
  return out;
}

// mask.rs:99-111
inline std::optional<size_t> nearest(const Grid& g, const Vec3& p) {
  const Vec3 c = g.point_to_index(p);
  size_t idx[3];
  for (int a = 0; a < 3; ++a) {
    const double r = std::floor(c[a] + 0.5);
    if (!(r >= 0.0 && r < double(g.size[a]))) return std::nullopt;
    idx[a] = size_t(r);
  }
  return g.linear_index(idx[0], idx[1], idx[2]);
}

// field_ops.rs:139-166: ITK VectorLinearInterpolateImageFunction (neighbours clamped).
inline Vec3 field_interp_index(const VectorField& f, const Vec3& c) {
  const auto& n = f.grid.size;
  size_t idx[3][2];
  double w[3][2];
  for (int a = 0; a < 3; ++a) {
    const double fl = std::floor(c[a]);
    const double d = c[a] - fl;
    const long long base = (long long)fl;
    const long long last = (long long)n[a] - 1;
    idx[a][0] = size_t(std::clamp(base, 0LL, last));
    idx[a][1] = size_t(std::clamp(base + 1, 0LL, last));
    w[a][0] = 1.0 - d;
    w[a][1] = d;
  }
  Vec3 out{0, 0, 0};
  for (int corner = 0; corner < 8; ++corner) {
    const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
    const double ov = w[0][bx] * w[1][by] * w[2][bz];
    if (ov != 0.0) {
      const Vec3& v = f.data[f.grid.linear_index(idx[0][bx], idx[1][by], idx[2][bz])];
      out[0] += ov * v[0];
      out[1] += ov * v[1];
      out[2] += ov * v[2];
    }
  }
  return out;
}

// field_ops.rs:170-175
inline std::optional<Vec3> field_interp(const VectorField& f, const Vec3& p) {
  const Vec3 c = f.grid.point_to_index(p);
  if (!f.grid.contains_continuous_index(c)) return std::nullopt;
  return field_interp_index(f, c);
}

// transform/field.rs:108-152: an f32 displacement-field transform p -> p + u(p), p outside.
struct DisplacementTransform {
  Grid grid;
  std::vector<float> data;  // 3 floats per voxel, interleaved
  Vec3 apply(const Vec3& p) const {
    const Vec3 c = grid.point_to_index(p);
    if (!grid.contains_continuous_index(c)) return p;
    long long base[3];
    double dist[3];
    for (int a = 0; a < 3; ++a) {
      const double f = std::floor(c[a]);
      base[a] = (long long)f;
      dist[a] = c[a] - f;
    }
    Vec3 out{0, 0, 0};
    for (int corner = 0; corner < 8; ++corner) {
      double ov = 1.0;
      size_t idx[3];
      for (int a = 0; a < 3; ++a) {
        const bool up = (corner >> a) & 1;
        const long long n = up ? base[a] + 1 : base[a];
        const double w = up ? dist[a] : 1.0 - dist[a];
        idx[a] = size_t(std::clamp(n, 0LL, (long long)grid.size[a] - 1));
        ov *= w;
      }
      if (ov != 0.0) {
        const size_t o = grid.linear_index(idx[0], idx[1], idx[2]) * 3;
        for (int k = 0; k < 3; ++k) out[k] += ov * double(data[o + k]);
      }
    }
    return {p[0] + out[0], p[1] + out[1], p[2] + out[2]};
  }
};

// A point map of the metric inputs: identity, an affine, or a displacement field.
struct PointMap {
  const AffineTransform* affine = nullptr;
  const DisplacementTransform* field = nullptr;
  Vec3 apply(const Vec3& p) const {
    if (field) return field->apply(p);
    if (affine) return affine->apply(p);
    return p;
  }
};

// ------------------------------------------------------------------- warps (cpu/mod.rs) --

// cpu/mod.rs:60-86
inline Image warp_through_field(const Image& src, const VectorField& field, const VectorField* pre, const AffineTransform* affine) {
  const Grid& g = field.grid;
  std::vector<double> out(g.num_voxels());
  parallel_for(out.size(), [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l) {
      auto [i, j, k] = g.ijk(l);
      const Vec3 x = g.index_to_point({double(i), double(j), double(k)});
      const Vec3& u = field.data[l];
      Vec3 p{x[0] + u[0], x[1] + u[1], x[2] + u[2]};
      if (pre) {
        if (auto d = field_interp(*pre, p)) p = {p[0] + (*d)[0], p[1] + (*d)[1], p[2] + (*d)[2]};
      }
      if (affine) p = affine->apply(p);
      auto v = lerp_point(src, p);
      out[l] = v ? *v : 0.0;
    }
  });
  return Image(g, std::move(out));
}

// cpu/mod.rs:89-132
inline Image warp_mask_through_field(const Image& mask, const VectorField& field, const VectorField* pre,
                                     const AffineTransform* affine, const Image* product) {
  const Grid& g = field.grid;
  if (product && product->grid != g) invalid("warp mask", "the product image is not on the grid of the field");
  std::vector<double> out(g.num_voxels());
  parallel_for(out.size(), [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l) {
      auto [i, j, k] = g.ijk(l);
      const Vec3 x = g.index_to_point({double(i), double(j), double(k)});
      const Vec3& u = field.data[l];
      Vec3 p{x[0] + u[0], x[1] + u[1], x[2] + u[2]};
      if (pre) {
        if (auto d = field_interp(*pre, p)) p = {p[0] + (*d)[0], p[1] + (*d)[1], p[2] + (*d)[2]};
      }
      if (affine) p = affine->apply(p);
      auto v = nearest(mask.grid, p);
      out[l] = (v && mask.data[*v] != 0.0) ? 1.0 : 0.0;
      if (product) out[l] *= product->data[l];
    }
  });
  return Image(g, std::move(out));
}

// ------------------------------------------------------------------- CC (metric/cc.rs) --

constexpr const char* kNoVoxelInMasks =
    "no voxel of the registration level lies inside both the fixed and the moving mask";

struct CcEvaluation {
  double value = 0.0;
  size_t num_valid = 0;
  std::vector<uint8_t> valid;
  std::vector<double> local_cc;
  std::vector<double> coefficient;
  std::vector<Vec3> derivative;
};

// metric/cc.rs:124-145: box sum of radius r along one axis, zero outside.
inline std::vector<std::array<double, 6>> box_sum_axis(const std::vector<std::array<double, 6>>& data,
                                                       const std::array<size_t, 3>& size, int axis, size_t r) {
  const size_t stride = axis == 0 ? 1 : axis == 1 ? size[0] : size[0] * size[1];
  const size_t n = size[axis];
  std::vector<std::array<double, 6>> out(data.size());
  parallel_for(data.size(), [&](size_t b, size_t e) {
    for (size_t idx = b; idx < e; ++idx) {
      const size_t pos = (idx / stride) % n;
      const size_t lo = pos >= r ? pos - r : 0;
      const size_t hi = std::min(pos + r, n - 1);
      const size_t base = idx - pos * stride;
      std::array<double, 6> acc{};
      for (size_t t = lo; t <= hi; ++t) {
        const auto& v = data[base + t * stride];
        for (int c = 0; c < 6; ++c) acc[c] += v[c];
      }
      out[idx] = acc;
    }
  });
  return out;
}

// metric/cc.rs:170-321 `CcMetric::evaluate_masked` with a virtual grid, fixed and moving images,
// their point maps and an optional mask on the virtual grid.
inline CcEvaluation cc_evaluate(const Grid& grid, const Image& fixed, const Image& moving, const PointMap& fmap,
                                const PointMap& mmap, const Image* mask, size_t radius, bool want_derivative) {
  const size_t n = grid.num_voxels();
  if (mask && mask->grid != grid) invalid("cc metric", "the mask is not on the virtual grid");
  CcEvaluation ev;
  ev.valid.assign(n, 0);
  std::vector<double> f(n, 0.0), m(n, 0.0);
  std::vector<Vec3> grad(want_derivative ? n : 0, Vec3{0, 0, 0});
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t idx = b; idx < e; ++idx) {
      if (mask && mask->data[idx] == 0.0) continue;
      auto [i, j, k] = grid.ijk(idx);
      const Vec3 p = grid.index_to_point({double(i), double(j), double(k)});
      const Vec3 fp = fmap.apply(p);
      auto fv = lerp_point(fixed, fp);
      if (!fv) continue;
      const Vec3 mp = mmap.apply(p);
      auto mv = lerp_point(moving, mp);
      if (!mv) continue;
      ev.valid[idx] = 1;
      f[idx] = *fv;
      m[idx] = *mv;
      if (want_derivative) grad[idx] = central_difference(moving, mp);
    }
  });
  for (size_t idx = 0; idx < n; ++idx) ev.num_valid += ev.valid[idx];
  if (ev.num_valid == 0)
    invalid("cc metric", mask ? kNoVoxelInMasks : "no virtual voxel maps inside both image buffers");
  std::vector<std::array<double, 6>> sums(n);
  for (size_t idx = 0; idx < n; ++idx)
    sums[idx] = ev.valid[idx] ? std::array<double, 6>{f[idx] * f[idx], m[idx] * m[idx], f[idx], m[idx], f[idx] * m[idx], 1.0}
                              : std::array<double, 6>{};
  for (int axis = 0; axis < 3; ++axis) sums = box_sum_axis(sums, grid.size, axis, radius);
  ev.local_cc.assign(n, 0.0);
  ev.coefficient.assign(n, 0.0);
  if (want_derivative) ev.derivative.assign(n, Vec3{0, 0, 0});
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t idx = b; idx < e; ++idx) {
      if (!ev.valid[idx]) continue;
      const auto& s = sums[idx];
      const double sum_f2 = s[0], sum_m2 = s[1], sum_f = s[2], sum_m = s[3], sum_fm = s[4], count = s[5];
      const double fmean = sum_f / count, mmean = sum_m / count;
      const double s_ff = sum_f2 - fmean * sum_f - fmean * sum_f + count * fmean * fmean;
      const double s_mm = sum_m2 - mmean * sum_m - mmean * sum_m + count * mmean * mmean;
      const double s_fm = sum_fm - mmean * sum_f - fmean * sum_m + count * mmean * fmean;
      const double fi = f[idx] - fmean, mi = m[idx] - mmean;
      const double product = s_ff * s_mm;
      double cc = 1.0;
      if (std::fabs(product) > kEps) cc = s_fm * s_fm / product;
      ev.local_cc[idx] = cc;
      if (want_derivative && s_ff > kEps && s_mm > kEps) {
        const double coeff = 2.0 * s_fm / product * (fi - s_fm / s_mm * mi);
        ev.coefficient[idx] = coeff;
        ev.derivative[idx] = {coeff * grad[idx][0], coeff * grad[idx][1], coeff * grad[idx][2]};
      }
    }
  });
  double total = 0.0;
  for (size_t idx = 0; idx < n; ++idx)
    if (ev.valid[idx]) total -= ev.local_cc[idx];
  ev.value = total / double(ev.num_valid);
  return ev;
}

// ------------------------------------------------------------ Mattes MI (metric/mi.rs) --

constexpr size_t kPadding = 2;

// mi.rs:62-74
inline double cubic_bspline(double u) {
  const double a = std::fabs(u);
  if (a < 1.0) {
    const double s = a * a;
    return (4.0 - 6.0 * s + 3.0 * s * a) / 6.0;
  }
  if (a < 2.0) {
    const double s = a * a;
    return (8.0 - 12.0 * a + 6.0 * s - s * a) / 6.0;
  }
  return 0.0;
}

// mi.rs:77-88
inline double cubic_bspline_derivative(double u) {
  if (u >= 0.0 && u < 1.0) return -2.0 * u + 1.5 * u * u;
  if (u > -1.0 && u < 0.0) return -2.0 * u - 1.5 * u * u;
  if (u >= 1.0 && u < 2.0) return -2.0 + 2.0 * u - 0.5 * u * u;
  if (u > -2.0 && u <= -1.0) return 2.0 + 2.0 * u + 0.5 * u * u;
  return 0.0;
}

// mi.rs:121-317 `MiBinning`
struct Binning {
  size_t bins = 32;
  double fixed_min = 0, fixed_max = 0, moving_min = 0, moving_max = 0;
  double fixed_bin_size = 0, moving_bin_size = 0, fixed_normalized_min = 0, moving_normalized_min = 0;

  static Binning from_ranges(size_t bins, double fmin, double fmax, double mmin, double mmax) {
    if (bins <= 2 * kPadding) invalid("mattes", "at least five histogram bins are required");
    if (fmax - fmin <= kEps) invalid("mattes", "the fixed image is constant");
    if (mmax - mmin <= kEps) invalid("mattes", "the moving image is constant");
    Binning b;
    b.bins = bins;
    b.fixed_min = fmin;
    b.fixed_max = fmax;
    b.moving_min = mmin;
    b.moving_max = mmax;
    const double span = double(bins - 2 * kPadding);
    b.fixed_bin_size = (fmax - fmin) / span;
    b.moving_bin_size = (mmax - mmin) / span;
    b.fixed_normalized_min = fmin / b.fixed_bin_size - double(kPadding);
    b.moving_normalized_min = mmin / b.moving_bin_size - double(kPadding);
    return b;
  }
  double moving_term(double v) const { return v / moving_bin_size - moving_normalized_min; }
  size_t clamp_bin(double term) const {
    // saturating truncation toward zero, as Rust's `as i64`
    long long idx = std::isnan(term) ? 0 : (long long)std::max(-1e18, std::min(1e18, term));
    return size_t(std::clamp(idx, (long long)kPadding, (long long)bins - 3));
  }
  size_t fixed_bin(double v) const { return clamp_bin(v / fixed_bin_size - fixed_normalized_min); }
};

template <typename Vec>
inline bool finite_range(const Vec& v, const std::vector<uint8_t>* keep, double& lo, double& hi) {
  bool any = false;
  for (size_t i = 0; i < v.size(); ++i) {
    if (keep && !(*keep)[i]) continue;
    if (!std::isfinite(v[i])) continue;
    if (!any) {
      lo = hi = v[i];
      any = true;
    } else {
      lo = std::min(lo, v[i]);
      hi = std::max(hi, v[i]);
    }
  }
  return any;
}

inline bool mask_inside(const Image& mask, const Vec3& p) {
  auto i = nearest(mask.grid, p);
  return i && mask.data[*i] != 0.0;
}

// mi.rs:171-184 `masked_range`
inline bool masked_range(const Image& image, const Image* mask, double& lo, double& hi) {
  if (!mask) return finite_range(image.data, nullptr, lo, hi);
  std::vector<uint8_t> keep(image.data.size());
  for (size_t l = 0; l < keep.size(); ++l) {
    auto [i, j, k] = image.grid.ijk(l);
    keep[l] = mask_inside(*mask, image.grid.index_to_point({double(i), double(j), double(k)}));
  }
  return finite_range(image.data, &keep, lo, hi);
}

// mi.rs:290-304 `MiBinning::dense`
inline Binning dense_binning(const Image& fixed, const Image& moving, const Image* fmask, const Image* mmask, size_t bins) {
  double flo, fhi, mlo, mhi;
  if (!masked_range(fixed, fmask, flo, fhi))
    invalid("mattes", fmask ? "no finite fixed-image voxel lies inside the fixed mask" : "the fixed image has no finite voxel");
  if (!masked_range(moving, mmask, mlo, mhi))
    invalid("mattes", mmask ? "no finite moving-image voxel lies inside the moving mask" : "the moving image has no finite voxel");
  return Binning::from_ranges(bins, flo, fhi, mlo, mhi);
}

struct Kahan {
  double sum = 0.0, comp = 0.0;
  void add(double x) {
    const double y = x - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
};

struct HistogramResult {
  double value = 0.0;
  std::vector<double> ratio, joint_pdf, fixed_marginal, moving_marginal;
};

// mi.rs:702-769 `finish_histogram`
inline HistogramResult finish_histogram(size_t bins, double moving_bin_size, size_t num_valid, std::vector<double> joint,
                                        std::vector<double> fixed_marginal, bool want_ratio) {
  if (num_valid == 0) invalid("mattes", "all samples map outside the moving image buffer");
  Kahan pdf_sum;
  for (double v : joint) pdf_sum.add(v);
  if (pdf_sum.sum < kEps) invalid("mattes", "joint PDF summed to zero");
  const double norm = 1.0 / pdf_sum.sum;
  for (double& v : joint) v *= norm;
  for (double& v : fixed_marginal) v *= norm;
  std::vector<double> mm(bins, 0.0);
  for (size_t col = 0; col < bins; ++col)
    for (size_t row = 0; row < bins; ++row) mm[col] += joint[row * bins + col];
  const double n_factor = -1.0 / (moving_bin_size * double(num_valid));
  HistogramResult r;
  if (want_ratio) r.ratio.assign(bins * bins, 0.0);
  double sum = 0.0;
  for (size_t f = 0; f < bins; ++f) {
    const double fm = fixed_marginal[f];
    if (fm <= kEps) continue;
    const double log_fm = std::log(fm);
    for (size_t m = 0; m < bins; ++m) {
      const double jv = joint[f * bins + m];
      if (mm[m] > kEps && jv > kEps) {
        const double pr = std::log(jv / mm[m]);
        sum += jv * (pr - log_fm);
        if (want_ratio) r.ratio[f * bins + m] = n_factor * pr;
      }
    }
  }
  r.value = -sum;
  r.joint_pdf = std::move(joint);
  r.fixed_marginal = std::move(fixed_marginal);
  r.moving_marginal = std::move(mm);
  return r;
}

struct MiDenseEvaluation {
  double value = 0.0;
  size_t num_valid = 0;
  Binning binning;
  std::vector<Vec3> derivative;
};

// mi.rs:578-676 `evaluate_dense`, with `dense_point` (mi.rs:550-576).
inline MiDenseEvaluation mi_dense(const Grid& grid, const Image& fixed, const Image& moving, const PointMap& fmap,
                                  const PointMap& mmap, const Image* fmask, const Image* mmask, size_t bins, bool want_derivative) {
  MiDenseEvaluation ev;
  ev.binning = dense_binning(fixed, moving, fmask, mmask, bins);
  const Binning& bn = ev.binning;
  const size_t n = grid.num_voxels();
  struct Pt {
    bool ok = false;
    size_t fbin = 0, mbin = 0;
    double term = 0.0;
    Vec3 g{0, 0, 0};
  };
  auto point = [&](size_t idx, bool deriv) {
    Pt pt;
    auto [i, j, k] = grid.ijk(idx);
    const Vec3 p = grid.index_to_point({double(i), double(j), double(k)});
    const Vec3 fp = fmap.apply(p);
    if (fmask && !mask_inside(*fmask, fp)) return pt;
    auto fv = lerp_point(fixed, fp);
    if (!fv) return pt;
    const Vec3 mp = mmap.apply(p);
    if (mmask && !mask_inside(*mmask, mp)) return pt;
    auto mv = lerp_point(moving, mp);
    if (!mv) return pt;
    if (!std::isfinite(*fv) || !std::isfinite(*mv) || *mv < bn.moving_min || *mv > bn.moving_max) return pt;
    pt.ok = true;
    pt.term = bn.moving_term(*mv);
    if (deriv) {
      pt.g = central_difference(moving, mp);
      if (!(std::isfinite(pt.g[0]) && std::isfinite(pt.g[1]) && std::isfinite(pt.g[2]))) pt.g = {0, 0, 0};
    }
    pt.fbin = bn.fixed_bin(*fv);
    pt.mbin = bn.clamp_bin(pt.term);
    return pt;
  };
  constexpr size_t kChunk = 8192;
  const size_t chunks = (n + kChunk - 1) / kChunk;
  std::vector<std::vector<double>> pj(chunks), pf(chunks);
  std::vector<size_t> pv(chunks, 0);
  parallel_chunks(n, kChunk, [&](size_t c, size_t b, size_t e) {
    pj[c].assign(bins * bins, 0.0);
    pf[c].assign(bins, 0.0);
    for (size_t idx = b; idx < e; ++idx) {
      Pt p = point(idx, false);
      if (!p.ok) continue;
      pf[c][p.fbin] += 1.0;
      for (size_t pm = p.mbin - 1; pm <= p.mbin + 2; ++pm) pj[c][p.fbin * bins + pm] += cubic_bspline(double(pm) - p.term);
      pv[c] += 1;
    }
  });
  std::vector<double> joint(bins * bins, 0.0), fmarg(bins, 0.0);
  for (size_t c = 0; c < chunks; ++c) {
    for (size_t q = 0; q < joint.size(); ++q) joint[q] += pj[c][q];
    for (size_t q = 0; q < bins; ++q) fmarg[q] += pf[c][q];
    ev.num_valid += pv[c];
  }
  if (ev.num_valid == 0 && (fmask || mmask)) invalid("mattes", kNoVoxelInMasks);
  HistogramResult fin = finish_histogram(bins, bn.moving_bin_size, ev.num_valid, std::move(joint), std::move(fmarg), want_derivative);
  ev.value = fin.value;
  if (want_derivative) {
    ev.derivative.assign(n, Vec3{0, 0, 0});
    parallel_for(n, [&](size_t b, size_t e) {
      for (size_t idx = b; idx < e; ++idx) {
        Pt p = point(idx, true);
        if (!p.ok) continue;
        double s = 0.0;
        for (size_t pm = p.mbin - 1; pm <= p.mbin + 2; ++pm)
          s += cubic_bspline_derivative(double(pm) - p.term) * fin.ratio[p.fbin * bins + pm];
        ev.derivative[idx] = {s * p.g[0], s * p.g[1], s * p.g[2]};
      }
    });
  }
  return ev;
}

struct MiSampledEvaluation {
  double value = 0.0;
  std::vector<double> derivative;
  size_t num_valid = 0;
  HistogramResult hist;
};

// mi.rs:364-511 `accumulate` + `MattesMi::evaluate` with prepared fixed bins (linear stage).
inline MiSampledEvaluation mi_sampled(const Binning& bn, const Image& moving, const std::vector<Vec3>& points,
                                      const std::vector<uint32_t>& fixed_bin, const LinearMovingTransform& t, bool want_derivative) {
  const size_t bins = bn.bins;
  const size_t params = t.optimised.count();
  const AffineTransform opt = t.optimised.to_affine();
  constexpr size_t kChunk = 8192;
  const size_t chunks = (points.size() + kChunk - 1) / kChunk;
  struct Partial {
    std::vector<double> joint, fm, dpdf;
    size_t valid = 0;
  };
  std::vector<Partial> parts(chunks);
  parallel_chunks(points.size(), kChunk, [&](size_t c, size_t b, size_t e) {
    Partial& P = parts[c];
    P.joint.assign(bins * bins, 0.0);
    P.fm.assign(bins, 0.0);
    if (want_derivative) P.dpdf.assign(bins * bins * params, 0.0);
    std::array<Vec3, 12> jac;
    for (size_t i = b; i < e; ++i) {
      const size_t f_index = fixed_bin[i];
      const Vec3 mp = t.outer.apply(opt.apply(points[i]));
      auto mv = lerp_point(moving, mp);
      if (!mv) continue;
      if (!std::isfinite(*mv) || *mv < bn.moving_min || *mv > bn.moving_max) continue;
      const double term = *mv / bn.moving_bin_size - bn.moving_normalized_min;
      const size_t m_index = bn.clamp_bin(term);
      P.fm[f_index] += 1.0;
      Vec3 grad{0, 0, 0};
      size_t ncols = 0;
      if (want_derivative) {
        grad = central_difference(moving, mp);
        ncols = t.jacobian(points[i], jac);
      }
      if (!(std::isfinite(grad[0]) && std::isfinite(grad[1]) && std::isfinite(grad[2]))) grad = {0, 0, 0};
      for (size_t pm = m_index - 1; pm <= m_index + 2; ++pm) {
        const double arg = double(pm) - term;
        const size_t cell = f_index * bins + pm;
        P.joint[cell] += cubic_bspline(arg);
        if (want_derivative) {
          const double db = cubic_bspline_derivative(arg);
          for (size_t mu = 0; mu < ncols; ++mu) {
            const double inner = jac[mu][0] * grad[0] + jac[mu][1] * grad[1] + jac[mu][2] * grad[2];
            P.dpdf[cell * params + mu] += inner * db;
          }
        }
      }
      P.valid += 1;
    }
  });
  std::vector<double> joint(bins * bins, 0.0), fm(bins, 0.0), dpdf(want_derivative ? bins * bins * params : 0, 0.0);
  MiSampledEvaluation ev;
  for (auto& P : parts) {
    for (size_t q = 0; q < joint.size(); ++q) joint[q] += P.joint[q];
    for (size_t q = 0; q < bins; ++q) fm[q] += P.fm[q];
    for (size_t q = 0; q < dpdf.size(); ++q) dpdf[q] += P.dpdf[q];
    ev.num_valid += P.valid;
  }
  ev.hist = finish_histogram(bins, bn.moving_bin_size, ev.num_valid, std::move(joint), std::move(fm), want_derivative);
  ev.value = ev.hist.value;
  if (want_derivative) {
    ev.derivative.assign(params, 0.0);
    for (size_t cell = 0; cell < ev.hist.ratio.size(); ++cell) {
      const double r = ev.hist.ratio[cell];
      if (r != 0.0)
        for (size_t mu = 0; mu < params; ++mu) ev.derivative[mu] += dpdf[cell * params + mu] * r;
    }
  }
  return ev;
}

// ------------------------------------------------------------- field ops (field_ops.rs) --

// field_ops.rs:395-417
inline double bessel_i0(double y) {
  const double d = std::fabs(y);
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1 + m * 0.45813e-2)))));
  }
  const double m = 3.75 / d;
  return (std::exp(d) / std::sqrt(d)) *
         (0.39894228 + m * (0.1328592e-1 + m * (0.225319e-2 + m * (-0.157565e-2 + m * (0.916281e-2 + m * (-0.2057706e-1 + m * (0.2635537e-1 + m * (-0.1647633e-1 + m * 0.392377e-2))))))));
}

// field_ops.rs:420-440
inline double bessel_i1(double y) {
  const double d = std::fabs(y);
  double acc;
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    acc = d * (0.5 + m * (0.87890594 + m * (0.51498869 + m * (0.15084934 + m * (0.2658733e-1 + m * (0.301532e-2 + m * 0.32411e-3))))));
  } else {
    const double m = 3.75 / d;
    double a = 0.2282967e-1 + m * (-0.2895312e-1 + m * (0.1787654e-1 - m * 0.420059e-2));
    a = 0.39894228 + m * (-0.3988024e-1 + m * (-0.362018e-2 + m * (0.163801e-2 + m * (-0.1031555e-1 + m * a))));
    acc = a * (std::exp(d) / std::sqrt(d));
  }
  return y < 0.0 ? -acc : acc;
}

// field_ops.rs:444-475 (Miller's downward recurrence)
inline double bessel_i(int n, double y) {
  const double kAccuracy = 40.0;
  if (y == 0.0) return 0.0;
  const double toy = 2.0 / std::fabs(y);
  double qip = 0.0, acc = 0.0, qi = 1.0;
  const int start = 2 * (n + int(std::sqrt(kAccuracy * double(n))));
  for (int j = start; j > 0; --j) {
    const double qim = qip + double(j) * toy * qi;
    qip = qi;
    qi = qim;
    if (std::fabs(qi) > 1.0e10) {
      acc *= 1.0e-10;
      qi *= 1.0e-10;
      qip *= 1.0e-10;
    }
    if (j == n) acc = qip;
  }
  acc *= bessel_i0(y) / qi;
  return (y < 0.0 && (n & 1)) ? -acc : acc;
}

// field_ops.rs:481-506 `gaussian_operator_coefficients`
inline std::vector<double> gaussian_taps(double variance, double max_error, size_t max_width) {
  const double et = std::exp(-variance);
  const double cap = 1.0 - max_error;
  std::vector<double> half{et * bessel_i0(variance)};
  double sum = half[0];
  half.push_back(et * bessel_i1(variance));
  sum += half[1] * 2.0;
  for (size_t i = 2; sum < cap; ++i) {
    half.push_back(et * bessel_i(int(i), variance));
    sum += half[i] * 2.0;
    if (half[i] <= 0.0) break;
    if (half.size() > max_width) break;
  }
  for (double& c : half) c /= sum;
  std::vector<double> full(half.rbegin(), half.rend() - 1);
  full.insert(full.end(), half.begin(), half.end());
  return full;
}

// field_ops.rs:508-533
inline std::vector<Vec3> convolve_axis(const std::vector<Vec3>& data, const std::array<size_t, 3>& size, int axis,
                                       const std::vector<double>& k) {
  const long long radius = (long long)(k.size() / 2);
  const size_t stride = axis == 0 ? 1 : axis == 1 ? size[0] : size[0] * size[1];
  const long long n = (long long)size[axis];
  std::vector<Vec3> out(data.size());
  parallel_for(data.size(), [&](size_t b, size_t e) {
    for (size_t idx = b; idx < e; ++idx) {
      const long long pos = (long long)((idx / stride) % size[axis]);
      const size_t base = idx - size_t(pos) * stride;
      Vec3 acc{0, 0, 0};
      for (size_t t = 0; t < k.size(); ++t) {
        const size_t q = size_t(std::clamp(pos + (long long)t - radius, 0LL, n - 1));
        const Vec3& v = data[base + q * stride];
        acc[0] += k[t] * v[0];
        acc[1] += k[t] * v[1];
        acc[2] += k[t] * v[2];
      }
      out[idx] = acc;
    }
  });
  return out;
}

// field_ops.rs:536-574 `gaussian_smooth_displacement_field`
inline VectorField smooth_field(const VectorField& field, double variance) {
  if (variance <= 0.0) return field;
  const auto& size = field.grid.size;
  std::vector<Vec3> cur = field.data;
  for (int axis = 0; axis < 3; ++axis) cur = convolve_axis(cur, size, axis, gaussian_taps(variance, 0.001, size[axis]));
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
  const Grid& g = field.grid;
  parallel_for(cur.size(), [&](size_t b, size_t e) {
    for (size_t idx = b; idx < e; ++idx) {
      auto [i, j, k] = g.ijk(idx);
      if (i == 0 || j == 0 || k == 0 || i == size[0] - 1 || j == size[1] - 1 || k == size[2] - 1) {
        cur[idx] = {0, 0, 0};
      } else {
        const Vec3& f = field.data[idx];
        cur[idx] = {cur[idx][0] * w1 + f[0] * w2, cur[idx][1] * w1 + f[1] * w2, cur[idx][2] * w1 + f[2] * w2};
      }
    }
  });
  return VectorField(g, std::move(cur));
}

inline double scaled_norm(const Vec3& v, const Vec3& s) {
  const double a = v[0] / s[0], b = v[1] / s[1], c = v[2] / s[2];
  return std::sqrt(a * a + b * b + c * c);
}

// field_ops.rs:191-204 (NaN propagates)
inline double max_scaled_norm(const VectorField& f) {
  double m = 0.0;
  for (const Vec3& v : f.data) {
    const double s = scaled_norm(v, f.grid.spacing);
    if (std::isnan(s) || std::isnan(m)) m = NAN;
    else m = std::max(m, s);
  }
  return m;
}

// field_ops.rs:207-213
inline double l2_norm(const VectorField& f) {
  std::vector<double> sq(f.data.size());
  for (size_t i = 0; i < sq.size(); ++i) sq[i] = f.data[i][0] * f.data[i][0] + f.data[i][1] * f.data[i][1] + f.data[i][2] * f.data[i][2];
  return std::sqrt(ordered_sum(sq));
}

// field_ops.rs:251-272 `compose_displacement_fields`
inline VectorField compose(const VectorField& displacement, const VectorField& warping) {
  if (displacement.grid != warping.grid) invalid("compose", "displacement and warping grids differ");
  const Grid& g = warping.grid;
  std::vector<Vec3> out(g.num_voxels());
  parallel_for(out.size(), [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l) {
      auto [i, j, k] = g.ijk(l);
      const Vec3 p1 = g.index_to_point({double(i), double(j), double(k)});
      const Vec3& w = warping.data[l];
      auto d = field_interp(displacement, {p1[0] + w[0], p1[1] + w[1], p1[2] + w[2]});
      const Vec3 dd = d ? *d : Vec3{0, 0, 0};
      out[l] = {w[0] + dd[0], w[1] + dd[1], w[2] + dd[2]};
    }
  });
  return VectorField(g, std::move(out));
}

struct InvertParams {
  uint32_t max_iterations = 20;
  double mean_error_tolerance = 0.001;
  double max_error_tolerance = 0.1;
  bool enforce_boundary_condition = true;
};

struct InvertStats {
  uint32_t iterations = 0;
  double mean_error_norm = 0.0;
  double max_error_norm = 0.0;
};

// A stop decision of the inversion loop that a scripted run took against the loop's own rule:
// after `body` bodies the rule would have continued (`natural`) or stopped, on these errors.
struct InvertDisagreement {
  uint32_t body = 0;
  bool natural = false;
  double max_error = 0.0, mean_error = 0.0;
};

// field_ops.rs:315-392 `invert_displacement_field`. With `forced_steps` >= 0 the loop runs exactly
// that many bodies instead of deciding by the tolerances (the lockstep comparison of a whole run
// replays another engine's step counts); each check where the tolerance rule would have decided
// otherwise is appended to `disagreements`.
inline std::pair<VectorField, InvertStats> invert(const VectorField& field, const VectorField* initial, const InvertParams& prm,
                                                  int forced_steps = -1,
                                                  std::vector<InvertDisagreement>* disagreements = nullptr) {
  const Grid& g = field.grid;
  VectorField inv = initial ? *initial : VectorField::zeros(g);
  if (initial && initial->grid != g) invalid("invert", "initial estimate grid differs");
  const double count = double(g.num_voxels());
  double max_error = std::numeric_limits<double>::max(), mean_error = std::numeric_limits<double>::max();
  uint32_t iteration = 0, executed = 0;
  const auto& n = g.size;
  for (;;) {
    const uint32_t current = iteration++;
    const bool natural = current < prm.max_iterations && max_error > prm.max_error_tolerance && mean_error > prm.mean_error_tolerance;
    const bool go = forced_steps < 0 ? natural : current < uint32_t(forced_steps);
    if (go != natural && disagreements) disagreements->push_back({current, natural, max_error, mean_error});
    if (!go) break;
    ++executed;
    VectorField composed = compose(field, inv);
    std::vector<double> norms(composed.data.size());
    for (size_t l = 0; l < norms.size(); ++l) norms[l] = scaled_norm(composed.data[l], g.spacing);
    mean_error = ordered_sum(norms) / count;
    max_error = 0.0;
    for (double v : norms) max_error = std::max(max_error, v);
    const double eps = current == 0 ? 0.75 : 0.5;
    const double limit = eps * max_error;
    parallel_for(norms.size(), [&](size_t b, size_t e) {
      for (size_t l = b; l < e; ++l) {
        const Vec3& er = composed.data[l];
        Vec3 up{-er[0], -er[1], -er[2]};
        if (norms[l] > limit) {
          const double f = limit / norms[l];
          up = {up[0] * f, up[1] * f, up[2] * f};
        }
        Vec3& psi = inv.data[l];
        psi = {psi[0] + up[0] * eps, psi[1] + up[1] * eps, psi[2] + up[2] * eps};
        if (prm.enforce_boundary_condition) {
          auto [i, j, k] = g.ijk(l);
          if (i == 0 || j == 0 || k == 0 || i == n[0] - 1 || j == n[1] - 1 || k == n[2] - 1) psi = {0, 0, 0};
        }
      }
    });
  }
  InvertStats st;
  st.iterations = executed;
  st.mean_error_norm = executed == 0 ? 0.0 : mean_error;
  st.max_error_norm = executed == 0 ? 0.0 : max_error;
  return {std::move(inv), st};
}

// field_ops.rs:580-605 `resample_field`
inline VectorField resample_field(const VectorField& field, const Grid& grid) {
  if (field.grid == grid) return field;
  const Grid& src = field.grid;
  std::vector<Image> comps;
  for (int c = 0; c < 3; ++c) {
    std::vector<double> d(field.data.size());
    for (size_t l = 0; l < d.size(); ++l) d[l] = field.data[l][c];
    comps.emplace_back(src, std::move(d));
  }
  std::vector<Vec3> out(grid.num_voxels());
  parallel_for(out.size(), [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l) {
      auto [i, j, k] = grid.ijk(l);
      const Vec3 p = grid.index_to_point({double(i), double(j), double(k)});
      const Vec3 c = src.point_to_index(p);
      if (!src.contains_continuous_index(c)) {
        out[l] = {0, 0, 0};
        continue;
      }
      out[l] = {lerp_index(comps[0], c), lerp_index(comps[1], c), lerp_index(comps[2], c)};
    }
  });
  return VectorField(grid, std::move(out));
}

// register.rs:82-142 `resample_isotropic` with `sample_linear_clamped`: the pipeline's isotropic
// grid (same origin and direction, floor((n - 1) * spacing / mm + 1e-9) + 1 voxels per axis);
// output voxel i samples source index min(i * mm / spacing, n - 1) trilinearly, corners outside
// the buffer or of zero weight skipped; the result is rounded to f32. A copy when every axis
// already has spacing mm (within 1e-6).
inline Image resample_isotropic(const Image& v, double mm) {
  if (!(std::isfinite(mm) && mm > 0.0)) invalid("register", "isotropic spacing must be positive");
  const Grid& g = v.grid;
  if (std::fabs(g.spacing[0] - mm) < 1e-6 && std::fabs(g.spacing[1] - mm) < 1e-6 && std::fabs(g.spacing[2] - mm) < 1e-6) return v;
  std::array<size_t, 3> out{1, 1, 1};
  std::array<std::vector<double>, 3> coords;
  for (int a = 0; a < 3; ++a) {
    const double last = double(g.size[a] - 1);
    const double ratio = mm / g.spacing[a];
    const size_t t = std::max<size_t>(size_t(std::floor(last * g.spacing[a] / mm + 1e-9)) + 1, 1);
    for (size_t i = 0; i < t; ++i) coords[a].push_back(std::min(double(i) * ratio, last));
    out[a] = t;
  }
  const Grid og(out, {mm, mm, mm}, g.origin, g.direction);
  std::vector<double> data(og.num_voxels());
  for (size_t k = 0; k < out[2]; ++k)
    for (size_t j = 0; j < out[1]; ++j)
      for (size_t i = 0; i < out[0]; ++i) {
        const double c[3] = {coords[0][i], coords[1][j], coords[2][k]};
        const double lo[3] = {std::floor(c[0]), std::floor(c[1]), std::floor(c[2])};
        double acc = 0.0;
        for (int corner = 0; corner < 8; ++corner) {
          size_t idx[3];
          double w = 1.0;
          bool inside = true;
          for (int a = 0; a < 3; ++a) {
            const int bit = (corner >> a) & 1;
            const long long q = (long long)lo[a] + bit;
            if (q < 0 || q >= (long long)g.size[a]) {
              inside = false;
              break;
            }
            idx[a] = size_t(q);
            const double fr = c[a] - lo[a];
            w *= bit ? fr : 1.0 - fr;
          }
          if (w == 0.0 || !inside) continue;
          acc += w * v.data[idx[0] + g.size[0] * (idx[1] + g.size[1] * idx[2])];
        }
        data[i + out[0] * (j + out[1] * k)] = double(float(acc));
      }
  return Image(og, std::move(data));
}

// ------------------------------------------------ level smoothing (pyramid.rs:184-465) --

// pyramid.rs:184-244: the zero-order coefficients of ITK's RecursiveGaussianImageFilter, as the
// reference forms them from Deriche's constants (n: causal numerator, d: denominator without its
// leading 1, m: anti-causal numerator, bn/bm: the boundary terms of constant extension).
struct DericheCoefs {
  double n[4], d[4], m[4], bn[4], bm[4];
};

inline DericheCoefs deriche_coefficients(double sigma, double spacing) {
  if (spacing < 1e-8) invalid("smoothing", "spacing is too small");
  const double sg = sigma / spacing;
  const double a1 = 1.3530, b1 = 1.8151, w1 = 0.6681, l1 = -1.3932;
  const double a2 = -0.3531, b2 = 0.0902, w2 = 2.0787, l2 = -1.3732;
  const double s1 = std::sin(w1 / sg), s2 = std::sin(w2 / sg);
  const double c1 = std::cos(w1 / sg), c2 = std::cos(w2 / sg);
  const double e1 = std::exp(l1 / sg), e2 = std::exp(l2 / sg);
  const double d1 = -2.0 * (e2 * c2 + e1 * c1);
  const double d2 = 4.0 * c2 * c1 * e1 * e2 + (e1 * e1 + e2 * e2);
  const double d3 = -2.0 * c1 * e1 * e2 * e2 + -2.0 * c2 * e2 * e1 * e1;
  const double d4 = e1 * e1 * e2 * e2;
  const double sd = 1.0 + d1 + d2 + d3 + d4;
  const double n0 = a1 + a2;
  const double n1 = e2 * (b2 * s2 - (a2 + 2.0 * a1) * c2) + e1 * (b1 * s1 - (a1 + 2.0 * a2) * c1);
  const double n2 = 2.0 * e1 * e2 * ((a1 + a2) * c2 * c1 - (b1 * c2 * s1 + b2 * c1 * s2)) + (a2 * e1 * e1 + a1 * e2 * e2);
  const double n3 = e2 * e1 * e1 * (b2 * s2 - a2 * c2) + e1 * e2 * e2 * (b1 * s1 - a1 * c1);
  const double alpha = 2.0 * (n0 + n1 + n2 + n3) / sd - n0;
  DericheCoefs k{};
  const double nn[4] = {n0 / alpha, n1 / alpha, n2 / alpha, n3 / alpha};
  const double dd[4] = {d1, d2, d3, d4};
  double sn = 0.0, sm = 0.0;
  for (int q = 0; q < 4; ++q) {
    k.n[q] = nn[q];
    k.d[q] = dd[q];
    k.m[q] = (q < 3 ? nn[q + 1] : 0.0) - dd[q] * nn[0];
    sn += k.n[q];
    sm += k.m[q];
  }
  for (int q = 0; q < 4; ++q) {
    k.bn[q] = dd[q] * sn / sd;
    k.bm[q] = dd[q] * sm / sd;
  }
  return k;
}

// pyramid.rs:246-304: the causal recursion y[i] = sum_q n_q x[i-q] - sum_q d_q y[i-1-q] and the
// anti-causal one a[i] = sum_q m_q x[i+1+q] - sum_q d_q a[i+1+q], added. The input is extended by
// its end values; the outputs before each recursion starts are replaced by v * bn_q (v * bm_q).
inline void deriche_filter_line(const DericheCoefs& k, const double* x, size_t len, double* y, double* a, double* out) {
  const long long L = (long long)len;
  const double v1 = x[0], v2 = x[len - 1];
  const auto at = [&](long long i) { return i < 0 ? v1 : (i >= L ? v2 : x[i]); };
  for (long long i = 0; i < L; ++i) {
    double fwd = 0.0, back = 0.0;
    for (int q = 0; q < 4; ++q) {
      fwd += k.n[q] * at(i - q);
      back += i - 1 - q >= 0 ? k.d[q] * y[i - 1 - q] : v1 * k.bn[q];
    }
    y[i] = fwd - back;
  }
  for (long long i = L - 1; i >= 0; --i) {
    double fwd = 0.0, back = 0.0;
    for (int q = 0; q < 4; ++q) {
      fwd += k.m[q] * at(i + 1 + q);
      back += i + 1 + q < L ? k.d[q] * a[i + 1 + q] : v2 * k.bm[q];
    }
    a[i] = fwd - back;
  }
  for (size_t i = 0; i < len; ++i) out[i] = y[i] + a[i];
}

// pyramid.rs:366-419: every line of `data` along `axis`, in place.
inline void deriche_axis(HostVec<double>& data, const std::array<size_t, 3>& size, int axis, const DericheCoefs& k) {
  const size_t nx = size[0], ny = size[1];
  const size_t len = size[axis];
  const size_t stride = axis == 0 ? 1 : (axis == 1 ? nx : nx * ny);
  const size_t lines = data.size() / len;
  parallel_chunks(lines, 64, [&](size_t, size_t b, size_t e) {
    std::vector<double> line(len), y(len), a(len), out(len);
    for (size_t l = b; l < e; ++l) {
      // the first voxel of line l: lines along x are rows, along y a row index within each plane,
      // along z a voxel of the first plane
      const size_t first = axis == 0 ? l * nx : (axis == 1 ? (l / nx) * nx * ny + l % nx : l);
      for (size_t t = 0; t < len; ++t) line[t] = data[first + t * stride];
      deriche_filter_line(k, line.data(), len, y.data(), a.data(), out.data());
      for (size_t t = 0; t < len; ++t) data[first + t * stride] = out[t];
    }
  });
}

// pyramid.rs:429-437: axes z, x, y, each with its own sigma in mm; every axis needs 4 voxels.
inline Image smooth_recursive_gaussian(const Image& vol, const Vec3& sigma_mm) {
  HostVec<double> out = vol.data;
  for (int axis : {2, 0, 1}) {
    if (vol.grid.size[axis] < 4) invalid("smoothing", "every axis needs at least 4 voxels");
    if (!(sigma_mm[axis] > 0.0)) invalid("smoothing", "sigma must be greater than zero");
    deriche_axis(out, vol.grid.size, axis, deriche_coefficients(sigma_mm[axis], vol.grid.spacing[axis]));
  }
  return Image(vol.grid, std::move(out));
}

// pyramid.rs:441-465 `sigma_mm_per_axis`, `smooth_for_level`: sigma in voxels scales by each
// axis's spacing; sigma <= 0 returns the image unchanged.
inline Image smooth_for_level(const Image& vol, double sigma, SigmaUnits units) {
  if (sigma <= 0.0) return vol;
  const Vec3& s = vol.grid.spacing;
  const Vec3 mm = units == SigmaUnits::Voxels ? Vec3{sigma * s[0], sigma * s[1], sigma * s[2]} : Vec3{sigma, sigma, sigma};
  return smooth_recursive_gaussian(vol, mm);
}

}  // namespace fw::oracle
