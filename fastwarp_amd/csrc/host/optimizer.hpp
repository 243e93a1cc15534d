// The optimiser of the linear antsRegistration stages (crates/fastwarp-reg/src/optimizer.rs):
// physical-shift parameter scales, ITK's conjugate gradient (Polak-Ribiere with the [0, 5] restart
// rule) and the golden-section line search over [lower, 1, upper] * learning rate. Host f64; the
// metric is any type with value(t) and value_and_derivative(t) (LinearMetric, optimizer.rs:60-65).
#pragma once

#include <atomic>
#include <chrono>

#include "convergence.hpp"
#include "grid.hpp"
#include "transform.hpp"

namespace fw {

// optimizer.rs:94-106 `CgSettings::ants`
struct CgSettings {
  double max_step_mm = 0.1;
  double lower_limit = 0.0;
  double upper_limit = 2.0;
  double epsilon = 0.2;
  uint32_t max_line_search_iterations = 20;
  uint32_t iterations = 0;
  double convergence_threshold = 1e-6;
  size_t convergence_window = 10;
  double small_parameter_variation = 0.01;
};

struct CgTrace {
  std::vector<double> values;
  bool converged = false;
  uint32_t iterations = 0;
  double best_value = std::numeric_limits<double>::max();
  std::vector<double> scales;
  double initial_learning_rate = 0.0, final_learning_rate = 0.0;
  uint64_t line_search_evaluations = 0;
  double iteration_seconds = 0.0;  // wall time of the iterations (evaluations and steps)
};

// optimizer.rs:136-160 `corner_points`
inline std::array<Vec3, 8> corner_points(const Grid& g) {
  std::array<Vec3, 8> out{};
  for (int i = 0; i < 8; ++i)
    out[i] = g.index_to_point({(i & 1) ? double(g.size[0] - 1) : 0.0, (i & 2) ? double(g.size[1] - 1) : 0.0,
                               (i & 4) ? double(g.size[2] - 1) : 0.0});
  return out;
}

inline double max_voxel_shift(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t, const std::vector<double>& delta) {
  std::vector<double> p = t.optimised.parameters();
  for (size_t i = 0; i < p.size(); ++i) p[i] += delta[i];
  const LinearMovingTransform moved = t.with_parameters(p);
  double mx = 0.0;
  for (const Vec3& c : corners) mx = std::max(mx, vnorm(vsub(moved.apply(c), t.apply(c))));
  return mx;
}

// optimizer.rs:185-215 `estimate_scales`
inline std::vector<double> estimate_scales(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t, double small) {
  const size_t n = t.optimised.count();
  std::vector<double> scales(n, 0.0);
  double min_nz = std::numeric_limits<double>::max();
  for (size_t i = 0; i < n; ++i) {
    std::vector<double> delta(n, 0.0);
    delta[i] = small;
    const double s = max_voxel_shift(corners, t, delta);
    scales[i] = s;
    if (s > kEps && s < min_nz) min_nz = s;
  }
  if (min_nz == std::numeric_limits<double>::max()) return std::vector<double>(n, 1.0);
  for (double& s : scales) {
    s = s <= kEps ? min_nz * min_nz : s * s;
    s *= 1.0 / (small * small);
  }
  return scales;
}

// optimizer.rs:219-231 `estimate_step_scale`
inline double estimate_step_scale(const std::array<Vec3, 8>& corners, const LinearMovingTransform& t,
                                  const std::vector<double>& step, double small) {
  double mx = 0.0;
  for (double v : step) mx = std::max(mx, std::fabs(v));
  if (mx <= kEps) return 0.0;
  const double factor = small / mx;
  std::vector<double> s(step.size());
  for (size_t i = 0; i < s.size(); ++i) s[i] = step[i] * factor;
  return max_voxel_shift(corners, t, s) / factor;
}

inline void check_cancel(const volatile int32_t* cancel) {
  if (cancel && *cancel) throw Cancelled();
}

// optimizer.rs:262-300 `LineSearch::golden` (RESPHI = 2 - 1.618034).
template <typename Metric>
double golden_search(Metric& metric, const LinearMovingTransform& start, const std::vector<double>& dir, const CgSettings& st,
                     double a, double b, double c, uint64_t& evaluations) {
  const double resphi = 2.0 - 1.618034;
  uint32_t iterations = 0;
  const std::vector<double> base = start.optimised.parameters();
  auto value_at = [&](double step) {
    ++evaluations;
    std::vector<double> p(base.size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = base[i] + step * dir[i];
    return metric.value(start.with_parameters(p));
  };
  double metric_b = std::numeric_limits<double>::max();
  for (;;) {
    if (iterations > st.max_line_search_iterations) return (c + a) / 2.0;
    ++iterations;
    const double x = (c - b > b - a) ? b + resphi * (c - b) : b - resphi * (b - a);
    if (std::fabs(c - a) < st.epsilon * (std::fabs(b) + std::fabs(x))) return (c + a) / 2.0;
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

// optimizer.rs:324-430 `optimise_level_async`: returns the best transform (ITK pairs the best
// value with the parameters after the step taken from it).
template <typename Metric>
std::pair<LinearMovingTransform, CgTrace> optimise_level(Metric& metric, const std::array<Vec3, 8>& corners,
                                                         LinearMovingTransform start, const CgSettings& st,
                                                         const volatile int32_t* cancel = nullptr) {
  const size_t n = start.optimised.count();
  const std::vector<double> scales = estimate_scales(corners, start, st.small_parameter_variation);
  LinearMovingTransform current = start, best = start;
  std::vector<double> conjugate(n, 0.0), last(n, 0.0);
  double lr = st.max_step_mm;
  WindowConvergenceMonitor monitor(st.convergence_window);
  double best_value = std::numeric_limits<double>::max();
  CgTrace trace;
  trace.scales = scales;
  trace.initial_learning_rate = trace.final_learning_rate = lr;
  uint32_t iteration = 0;
  const auto t_iters = std::chrono::steady_clock::now();
  while (iteration < st.iterations) {
    check_cancel(cancel);
    auto [value, gradient] = metric.value_and_derivative(current);
    monitor.add_energy_value(value);
    if (monitor.convergence_value() <= st.convergence_threshold) {
      trace.converged = true;
      break;
    }
    trace.values.push_back(value);
    for (size_t i = 0; i < n; ++i) gradient[i] *= 1.0 / scales[i];
    if (iteration == 0) {
      const double ss = estimate_step_scale(corners, current, gradient, st.small_parameter_variation);
      lr = ss <= kEps ? 1.0 : st.max_step_mm / ss;
      trace.initial_learning_rate = lr;
    }
    double gamma = 0.0;
    double denom = 0.0;
    for (size_t i = 0; i < n; ++i) denom += last[i] * last[i];
    if (denom > kEps) {
      double num = 0.0;
      for (size_t i = 0; i < n; ++i) num += (gradient[i] - last[i]) * gradient[i];
      gamma = num / denom;
    }
    if (!(gamma >= 0.0 && gamma <= 5.0)) gamma = 0.0;
    last = gradient;
    for (size_t i = 0; i < n; ++i) conjugate[i] = gradient[i] + conjugate[i] * gamma;
    lr = golden_search(metric, current, conjugate, st, lr * st.lower_limit, lr, lr * st.upper_limit, trace.line_search_evaluations);
    std::vector<double> p = current.optimised.parameters();
    for (size_t i = 0; i < n; ++i) p[i] += lr * conjugate[i];
    current = current.with_parameters(p);
    if (value < best_value) {
      best_value = value;
      best = current;
    }
    ++iteration;
  }
  trace.iterations = iteration;
  trace.iteration_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iters).count();
  trace.best_value = best_value;
  trace.final_learning_rate = lr;
  return {best, trace};
}

}  // namespace fw
