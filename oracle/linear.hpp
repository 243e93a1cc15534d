// ORACLE — test infrastructure only. The CPU engine's own linear chain: the Rigid and Affine stages
// (crates/fastwarp-reg/src/linear.rs:873-1068, 1148-1212) and ITK's conjugate-gradient line-search
// optimizer (optimizer.rs:185-430) restated directly over the oracle's Mattes MI, independent of the
// product's linear driver (fastwarp_amd/csrc/host/linear.hpp and optimizer.hpp, which prefetch
// samples, reuse levels between stages and search speculatively for the device). Shared with the
// product are only the pinned host pieces: MT19937 and the REGULAR sampler (published outputs),
// the transforms (Jacobians by finite differences), the bin geometry (bit for bit against the
// oracle's) and the window convergence monitor (closed form). The registration pipeline picks
// this chain for the CPU engine by argument-dependent lookup of linear_chain_of.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "engine.hpp"

namespace fw::oracle {

namespace linear_detail {

constexpr double kEps = 2.220446049250313e-16;
constexpr double kResphi = 2.0 - 1.618034;  // optimizer.rs RESPHI

// optimizer.rs:136-160 `corner_points`
inline std::array<Vec3, 8> corners_of(const Grid& g) {
  std::array<Vec3, 8> out{};
  for (int i = 0; i < 8; ++i) {
    const double x = (i & 1) ? double(g.size[0] - 1) : 0.0;
    const double y = (i & 2) ? double(g.size[1] - 1) : 0.0;
    const double z = (i & 4) ? double(g.size[2] - 1) : 0.0;
    out[i] = g.index_to_point({x, y, z});
  }
  return out;
}

// optimizer.rs:162-175 `max_voxel_shift`
inline double max_shift(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t, const std::vector<double>& delta) {
  std::vector<double> p = t.optimised.parameters();
  for (size_t i = 0; i < p.size(); ++i) p[i] += delta[i];
  const LinearMovingTransform moved = t.with_parameters(p);
  double m = 0.0;
  for (const Vec3& c : corners) {
    const Vec3 a = moved.apply(c), b = t.apply(c);
    const Vec3 d{a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    m = std::max(m, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
  }
  return m;
}

// optimizer.rs:177-208 `estimate_scales`
inline std::vector<double> scales_of(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t, double small) {
  const size_t n = t.optimised.count();
  std::vector<double> s(n, 0.0);
  double min_non_zero = std::numeric_limits<double>::max();
  for (size_t i = 0; i < n; ++i) {
    std::vector<double> delta(n, 0.0);
    delta[i] = small;
    s[i] = max_shift(corners, t, delta);
    if (s[i] > kEps && s[i] < min_non_zero) min_non_zero = s[i];
  }
  if (min_non_zero == std::numeric_limits<double>::max()) return std::vector<double>(n, 1.0);
  for (double& v : s) {
    if (v <= kEps) v = min_non_zero * min_non_zero;
    else v *= v;
    v *= 1.0 / (small * small);
  }
  return s;
}

// optimizer.rs:210-225 `estimate_step_scale`
inline double step_scale_of(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t, const std::vector<double>& step,
                            double small) {
  double max_step = 0.0;
  for (double v : step) max_step = std::max(max_step, std::fabs(v));
  if (max_step <= kEps) return 0.0;
  const double factor = small / max_step;
  std::vector<double> s(step.size());
  for (size_t i = 0; i < step.size(); ++i) s[i] = step[i] * factor;
  return max_shift(corners, t, s) / factor;
}

inline double inner(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// optimizer.rs:262-300 `LineSearch::golden` over steps along `direction` from `start`.
template <typename Level>
double golden(Level& metric, const LinearMovingTransform& start, const std::vector<double>& direction,
              uint32_t max_iterations, double epsilon, double a, double b, double c, uint64_t& evaluations) {
  const std::vector<double> p0 = start.optimised.parameters();
  auto value_at = [&](double step) {
    ++evaluations;
    std::vector<double> p(p0.size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = p0[i] + step * direction[i];
    return metric.value(start.with_parameters(p));
  };
  double metric_b = std::numeric_limits<double>::max();
  uint32_t iterations = 0;
  for (;;) {
    if (iterations > max_iterations) return (c + a) / 2.0;
    ++iterations;
    const double x = c - b > b - a ? b + kResphi * (c - b) : b - kResphi * (b - a);
    if (std::fabs(c - a) < epsilon * (std::fabs(b) + std::fabs(x))) return (c + a) / 2.0;
    const double metric_x = value_at(x);
    if (metric_b == std::numeric_limits<double>::max()) metric_b = value_at(b);
    if (metric_x < metric_b) {
      if (c - b > b - a) {
        a = b;
        b = x;
      } else {
        c = b;
        b = x;
      }
      metric_b = metric_x;
    } else if (c - b > b - a) {
      c = x;
    } else if (metric_x == std::numeric_limits<double>::max()) {
      c = b;
      b = x;
      metric_b = metric_x;
    } else {
      a = x;
    }
  }
}

// optimizer.rs:324-430 `optimise_level_async`, with CgSettings::ants (optimizer.rs:94-106).
template <typename Level>
std::pair<LinearMovingTransform, CgTrace> optimise(Level& metric, const std::array<Vec3, 8>& corners, LinearMovingTransform start,
                                                   double max_step_mm, uint32_t max_iterations, double threshold, size_t window,
                                                   const volatile int32_t* cancel) {
  constexpr double lower = 0.0, upper = 2.0, epsilon = 0.2, small = 0.01;
  constexpr uint32_t line_iterations = 20;
  const size_t n = start.optimised.count();
  const std::vector<double> scales = scales_of(corners, start, small);
  LinearMovingTransform current = start, best = start;
  std::vector<double> conjugate(n, 0.0), last(n, 0.0);
  double learning_rate = max_step_mm;
  WindowConvergenceMonitor monitor(window);
  double best_value = std::numeric_limits<double>::max();
  CgTrace trace;
  trace.scales = scales;
  trace.initial_learning_rate = trace.final_learning_rate = learning_rate;
  uint32_t iteration = 0;
  while (iteration < max_iterations) {
    check_cancel(cancel);
    auto [value, gradient] = metric.value_and_derivative(current);
    monitor.add_energy_value(value);
    if (monitor.convergence_value() <= threshold) {
      trace.converged = true;
      break;
    }
    trace.values.push_back(value);
    for (size_t i = 0; i < n; ++i) gradient[i] *= 1.0 / scales[i];
    if (iteration == 0) {
      const double s = step_scale_of(corners, current, gradient, small);
      learning_rate = s <= kEps ? 1.0 : max_step_mm / s;
      trace.initial_learning_rate = learning_rate;
    }
    double gamma = 0.0;
    const double denom = inner(last, last);
    if (denom > kEps) {
      std::vector<double> diff(n);
      for (size_t i = 0; i < n; ++i) diff[i] = gradient[i] - last[i];
      gamma = inner(diff, gradient) / denom;
    }
    if (!(gamma >= 0.0 && gamma <= 5.0)) gamma = 0.0;
    last = gradient;
    for (size_t i = 0; i < n; ++i) conjugate[i] = gradient[i] + conjugate[i] * gamma;
    learning_rate = golden(metric, current, conjugate, line_iterations, epsilon, learning_rate * lower, learning_rate,
                           learning_rate * upper, trace.line_search_evaluations);
    std::vector<double> p = current.optimised.parameters();
    for (size_t i = 0; i < n; ++i) p[i] = p[i] + learning_rate * conjugate[i];
    current = current.with_parameters(p);
    if (value < best_value) {
      best_value = value;
      best = current;
    }
    ++iteration;
  }
  trace.iterations = iteration;
  trace.best_value = best_value;
  trace.final_learning_rate = learning_rate;
  return {best, trace};
}

// linear.rs:873-1068 `run_linear_stage_cached_async` on the first channel of `cache`.
inline LinearStageResult stage(CpuLinearEngine& engine, const LinearRecipe& recipe, SmoothCache<CpuBackend>& cache,
                               const AffineTransform& outer, const ParametricTransform& start, const volatile int32_t* cancel) {
  recipe.validate();
  if (start.count() != (recipe.kind == LinearKind::Rigid ? 6u : 12u))
    invalid("linear stage", "start transform does not match the stage kind");
  const Grid fixed_grid = cache.fixed_grid(0);
  SeedSequence seeds(recipe.seed);
  LinearMovingTransform current{outer, start};
  LinearStageResult result;
  for (size_t level = 0; level < recipe.levels(); ++level) {
    check_cancel(cancel);
    const auto factors = shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing);
    const Grid grid = level_virtual_grid(fixed_grid, factors);
    const auto images = cache.get(0, recipe.smoothing_sigmas[level], recipe.sigma_units);
    const Image& fixed = *images.fixed;
    const Image& moving = *images.moving;
    const HostVec<Vec3> raw = metric_sample_points(grid, recipe.sampling_percentage, seeds);
    LinearLevelTrace tr;
    tr.level = level;
    tr.shrink_factors = factors;
    tr.grid_size = grid.size;
    // linear.rs:474-489 `degenerate_level` and mi.rs:193-207: the nearest-voxel values observed
    size_t observed = 0;
    double flo = 0.0, fhi = 0.0;
    for (const Vec3& p : raw) {
      size_t idx;
      if (!nearest_voxel(fixed.grid, p, idx)) continue;
      const double v = fixed.data[idx];
      if (!std::isfinite(v)) continue;
      flo = observed ? std::min(flo, v) : v;
      fhi = observed ? std::max(fhi, v) : v;
      ++observed;
    }
    if (observed == 0) tr.skipped = "no sample point lands on the fixed image";
    else if (fhi - flo <= kEps) tr.skipped = "the fixed image takes a single value at its sample points";
    if (tr.skipped.empty()) {
      // linear.rs:251-279 `prepare_samples`: the fixed range over the samples, the moving range
      // over the whole moving image, the fixed bin of every sample with a finite linear value
      double mlo = 0.0, mhi = 0.0;
      bool any = false;
      for (double v : moving.data) {
        if (!std::isfinite(v)) continue;
        mlo = any ? std::min(mlo, v) : v;
        mhi = any ? std::max(mhi, v) : v;
        any = true;
      }
      if (!any) invalid("mattes", "the moving image has no finite voxel");
      if (recipe.bins <= 4) invalid("mattes", "at least five histogram bins are required");
      const MiBinning binning = MiBinning::from_ranges(recipe.bins, flo, fhi, mlo, mhi);
      HostVec<Vec3> points;
      HostVec<uint32_t> bins;
      for (const Vec3& p : raw) {
        double v;
        if (!linear_at_point(fixed, p, v) || !std::isfinite(v)) continue;
        points.push_back(p);
        bins.push_back(uint32_t(binning.fixed_bin(v)));
      }
      if (points.empty()) {
        tr.skipped = "no sample point lies inside the fixed image";
      } else {
        tr.samples = points.size();
        CpuLinearEngine::Level metric = engine.prepare_level(images.moving, points, bins, binning);
        auto [next, cg] = optimise(metric, corners_of(grid), current, recipe.gradient_step, recipe.iterations[level],
                                   recipe.convergence_threshold, recipe.convergence_window, cancel);
        current = next;
        tr.cg = std::move(cg);
      }
    }
    tr.parameters = current.optimised.parameters();
    result.levels.push_back(std::move(tr));
  }
  bool all_skipped = !result.levels.empty();
  for (const auto& l : result.levels) all_skipped = all_skipped && !l.skipped.empty();
  if (all_skipped)
    invalid("linear stage", "every level of the stage has too few usable samples for Mattes mutual information");
  result.transform = current.optimised;
  result.affine = current.optimised.to_affine();
  return result;
}

}  // namespace linear_detail

// linear.rs:1148-1212 `run_linear_chain_cached_async` from a given initial transform.
inline LinearChainResult oracle_linear_chain(CpuLinearEngine& engine, const LinearRecipe& rigid, const LinearRecipe& affine,
                                             const AffineTransform& initial, SmoothCache<CpuBackend>& cache,
                                             const volatile int32_t* cancel) {
  if (rigid.kind != LinearKind::Rigid || affine.kind != LinearKind::Affine)
    invalid("linear chain", "expected a rigid then an affine recipe");
  LinearChainResult out;
  out.initial = initial;
  out.rigid = linear_detail::stage(engine, rigid, cache, initial, ParametricTransform::euler3d({0, 0, 0, 0, 0, 0}, initial.center),
                                   cancel);
  const AffineTransform after_rigid = AffineTransform::compose(initial, out.rigid.affine);
  const ParametricTransform start = ParametricTransform::affine(AffineTransform{kIdentity3, {0, 0, 0}, out.rigid.affine.center});
  out.affine = linear_detail::stage(engine, affine, cache, after_rigid, start, cancel);
  out.total = with_center(AffineTransform::compose(after_rigid, out.affine.affine), out.affine.affine.center);
  return out;
}

// Whether the CPU engine runs the product's generic linear chain over its ops instead of its own
// (a test compares the two).
inline bool& use_generic_linear() {
  static bool on = false;
  return on;
}

// The linear chain of the CPU engine's registrations (found by argument-dependent lookup from the
// registration pipeline in place of the generic chain).
inline LinearChainResult linear_chain_of(CpuLinearEngine& engine, const LinearRecipe& rigid, const LinearRecipe& affine,
                                         const AffineTransform& initial, SmoothCache<CpuBackend>& cache,
                                         const volatile int32_t* cancel) {
  if (use_generic_linear()) return fw::run_linear_chain(engine, rigid, affine, initial, cache, cancel);
  return oracle_linear_chain(engine, rigid, affine, initial, cache, cancel);
}

}  // namespace fw::oracle
