// ORACLE — test infrastructure only. The f64 CPU engine: the oracle's ops behind the backend and
// linear-engine interfaces that the generic drivers (SyN loop, linear stages) are written against,
// as the reference's `CpuBackend` and `CpuLinearEngine` (crates/fastwarp-reg/src/cpu/mod.rs:386-597,
// linear.rs:323-400) sit behind its `Backend` and `LinearEngine` traits.
#pragma once

#include "../fastwarp_amd/csrc/host/backend_types.hpp"
#include "../fastwarp_amd/csrc/host/register.hpp"
#include "oracle.hpp"

namespace fw::oracle {

// The lockstep comparison: the inversion step counts of another engine's run, replayed in call
// order, and the stop decisions of the oracle's own rule that differ from them.
struct InversionScript {
  std::vector<int32_t> steps;
  size_t next = 0;
  std::vector<InvertDisagreement> disagreements;
  std::vector<size_t> disagreement_calls;  // the call index of each disagreement
};

class CpuBackend {
 public:
  InversionScript* script = nullptr;

  using ImageH = Image;
  using FieldH = VectorField;

  void release_idle_buffers() {}
  Image upload_image(const Image& v) { return v; }
  Image smooth_image(const Image& v, double sigma, SigmaUnits units) { return smooth_for_level(v, sigma, units); }
  // mi.rs:161-207 and interp.rs:43-87: nearest-voxel and linear values, NaN outside.
  void sample_values(const Image& img, const HostVec<Vec3>& points, HostVec<double>& nearest,
                     HostVec<double>& linear) {
    nearest.assign(points.size(), std::nan(""));
    linear.assign(points.size(), std::nan(""));
    for (size_t i = 0; i < points.size(); ++i) {
      size_t idx;
      if (nearest_voxel(img.grid, points[i], idx)) nearest[i] = img.data[idx];
      double v;
      if (linear_at_point(img, points[i], v)) linear[i] = v;
    }
  }
  bool finite_range(const Image& img, double& lo, double& hi) {
    bool any = false;
    for (double v : img.data) {
      if (!std::isfinite(v)) continue;
      lo = any ? std::min(lo, v) : v;
      hi = any ? std::max(hi, v) : v;
      any = true;
    }
    return any;
  }
  VectorField zero_field(const Grid& g) { return VectorField::zeros(g); }
  VectorField download_field(const VectorField& f) { return f; }
  void download_field_to(const VectorField& f, double* aos) {
    for (size_t i = 0; i < f.data.size(); ++i)
      for (int c = 0; c < 3; ++c) aos[3 * i + c] = f.data[i][c];
  }
  Grid field_grid(const VectorField& f) const { return f.grid; }
  Image warp_through_field(const Image& src, const VectorField& f, const VectorField* pre, const AffineTransform* a) {
    return oracle::warp_through_field(src, f, pre, a);
  }
  Image warp_mask_through_field(const Image& mask, const VectorField& f, const VectorField* pre, const AffineTransform* a,
                                const Image* product) {
    return oracle::warp_mask_through_field(mask, f, pre, a, product);
  }
  MetricPair<VectorField> cc_pair(const Image& f, const Image& m, const Image* valid, size_t radius, bool norms) {
    MetricPair<VectorField> out;
    PointMap id;
    const Image* ab[2][2] = {{&f, &m}, {&m, &f}};
    for (int d = 0; d < 2; ++d) {
      CcEvaluation e = cc_evaluate(f.grid, *ab[d][0], *ab[d][1], id, id, valid, radius, true);
      out.values[d] = e.value;
      out.fields[d] = VectorField(f.grid, std::move(e.derivative));
      if (norms) out.norms[d] = l2_norm(out.fields[d]);
    }
    return out;
  }
  MetricPair<VectorField> mi_pair(const Image& f, const Image& m, const Image* fm, const Image* mm, size_t bins, bool norms) {
    MetricPair<VectorField> out;
    PointMap id;
    for (int d = 0; d < 2; ++d) {
      MiDenseEvaluation e = d == 0 ? mi_dense(f.grid, f, m, id, id, fm, mm, bins, true) : mi_dense(f.grid, m, f, id, id, mm, fm, bins, true);
      out.values[d] = e.value;
      out.fields[d] = VectorField(f.grid, std::move(e.derivative));
      if (norms) out.norms[d] = l2_norm(out.fields[d]);
    }
    return out;
  }
  void axpy(VectorField& acc, double a, const VectorField& x) {
    if (acc.grid != x.grid) invalid("field axpy", "grids differ");
    for (size_t i = 0; i < acc.data.size(); ++i)
      for (int c = 0; c < 3; ++c) acc.data[i][c] += a * x.data[i][c];
  }
  void scale(VectorField& f, double s) {
    for (auto& v : f.data)
      for (int c = 0; c < 3; ++c) v[c] *= s;
  }
  VectorField smooth_field(VectorField f, double variance) { return oracle::smooth_field(f, variance); }
  std::pair<double, double> max_scaled_norm_pair(const VectorField& a, const VectorField* b) {
    return {max_scaled_norm(a), b ? max_scaled_norm(*b) : 0.0};
  }
  // The deferred forms of the SyN driver: computed at once on the host.
  MetricPair<VectorField> cc_pair_deferred(const Image& f, const Image& m, size_t radius) {
    MetricPair<VectorField> out = cc_pair(f, m, nullptr, radius, false);
    out.later = Pending<std::array<double, 2>>(std::array<double, 2>{out.values[0], out.values[1]});
    return out;
  }
  Pending<std::pair<double, double>> scale_to_step(VectorField& a, VectorField& b, double step) {
    const auto m = max_scaled_norm_pair(a, &b);
    if (std::isfinite(m.first) && std::isfinite(m.second)) {
      scale(a, m.first > 0.0 ? step / m.first : step);
      scale(b, m.second > 0.0 ? step / m.second : step);
    }
    return Pending<std::pair<double, double>>(m);
  }
  std::pair<VectorField, Pending<fw::InvertStats>> invert_deferred(const VectorField& field, VectorField initial,
                                                                  const fw::InvertParams& p) {
    auto r = invert(field, std::move(initial), p);
    return {std::move(r.first), Pending<fw::InvertStats>(r.second)};
  }
  void settle() {}
  VectorField compose(const VectorField& u, const VectorField& w) { return oracle::compose(u, w); }
  std::pair<VectorField, fw::InvertStats> invert(const VectorField& field, VectorField initial, const fw::InvertParams& p) {
    oracle::InvertParams q{p.max_iterations, p.mean_error_tolerance, p.max_error_tolerance, p.enforce_boundary_condition};
    int forced = -1;
    std::vector<InvertDisagreement>* dis = nullptr;
    size_t before = 0, call = 0;
    if (script) {
      if (script->next >= script->steps.size()) invalid("lockstep", "the script has fewer inversions than the run");
      call = script->next;
      forced = script->steps[script->next++];
      dis = &script->disagreements;
      before = dis->size();
    }
    auto r = oracle::invert(field, &initial, q, forced, dis);
    if (script)
      for (size_t k = before; k < dis->size(); ++k) script->disagreement_calls.push_back(call);
    return {std::move(r.first), fw::InvertStats{r.second.iterations, r.second.mean_error_norm, r.second.max_error_norm}};
  }
  VectorField resample_field(const VectorField& f, const Grid& g) { return oracle::resample_field(f, g); }
};

class CpuLinearEngine {
 public:
  struct Level {
    Image moving;
    std::vector<Vec3> points;
    std::vector<uint32_t> fixed_bin;
    Binning binning;
    double value(const LinearMovingTransform& t) { return mi_sampled(binning, moving, points, fixed_bin, t, false).value; }
    std::pair<double, std::vector<double>> value_and_derivative(const LinearMovingTransform& t) {
      MiSampledEvaluation e = mi_sampled(binning, moving, points, fixed_bin, t, true);
      return {e.value, e.derivative};
    }
  };
  Level prepare_level(std::shared_ptr<const Image> moving, const HostVec<Vec3>& points,
                      const HostVec<uint32_t>& fixed_bin, const MiBinning& b) {
    Level l;
    l.moving = *moving;
    l.points.assign(points.begin(), points.end());
    l.fixed_bin.assign(fixed_bin.begin(), fixed_bin.end());
    l.binning = Binning::from_ranges(b.bins, b.fixed_min, b.fixed_max, b.moving_min, b.moving_max);
    return l;
  }
};

}  // namespace fw::oracle
