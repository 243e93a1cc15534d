// The optimiser of the linear antsRegistration stages (crates/fastwarp-reg/src/optimizer.rs):
// physical-shift parameter scales, ITK's conjugate gradient (Polak-Ribiere with the [0, 5] restart
// rule) and the golden-section line search over [lower, 1, upper] * learning rate. Host f64; the
// metric is any type with value(t) and value_and_derivative(t) (LinearMetric, optimizer.rs:60-65).
#pragma once

#include <atomic>
#include <chrono>
#include <limits>
#include <type_traits>
#include <utility>

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

// The flag is set from another thread: read it with an atomic load.
inline void check_cancel(const volatile int32_t* cancel) {
  if (cancel && __atomic_load_n(cancel, __ATOMIC_RELAXED)) throw Cancelled();
}

// optimizer.rs:262-300 `LineSearch::golden` (RESPHI = 2 - 1.618034), as two pure steps over the
// bracket so that the plain search and the speculative one below share them.
struct GoldenBracket {
  double a, b, c;
  double metric_b;  // the reference's sentinel f64::MAX while f(b) has not been evaluated
  uint32_t iterations;
};

// The next probe x of bracket g (and the bracket with the iteration counted), or false when the
// search ends there, with its return value in `result`.
inline bool golden_probe(const GoldenBracket& g, const CgSettings& st, double& x, double& result, GoldenBracket& next) {
  const double resphi = 2.0 - 1.618034;
  if (g.iterations > st.max_line_search_iterations) {
    result = (g.c + g.a) / 2.0;
    return false;
  }
  next = g;
  ++next.iterations;
  x = (g.c - g.b > g.b - g.a) ? g.b + resphi * (g.c - g.b) : g.b - resphi * (g.b - g.a);
  if (std::fabs(g.c - g.a) < st.epsilon * (std::fabs(g.b) + std::fabs(x))) {
    result = (g.c + g.a) / 2.0;
    return false;
  }
  return true;
}

// The bracket after probing x: f(x) = mx, f(b) = mb.
inline GoldenBracket golden_update(GoldenBracket g, double x, double mx, double mb) {
  constexpr double kMax = std::numeric_limits<double>::max();
  g.metric_b = mb;
  if (mx < mb) {
    if (g.c - g.b > g.b - g.a) {
      g.a = g.b;
      g.b = x;
    } else {
      g.c = g.b;
      g.b = x;
    }
    g.metric_b = mx;
  } else if (g.c - g.b > g.b - g.a) {
    g.c = x;
  } else if (mx == kMax) {
    g.c = g.b;
    g.b = x;
    g.metric_b = mx;
  } else {
    g.a = x;
  }
  return g;
}

template <typename Metric>
double golden_search(Metric& metric, const LinearMovingTransform& start, const std::vector<double>& dir, const CgSettings& st,
                     double a, double b, double c, uint64_t& evaluations) {
  constexpr double kMax = std::numeric_limits<double>::max();
  const std::vector<double> base = start.optimised.parameters();
  auto value_at = [&](double step) {
    ++evaluations;
    std::vector<double> p(base.size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = base[i] + step * dir[i];
    return metric.value(start.with_parameters(p));
  };
  GoldenBracket g{a, b, c, kMax, 0};
  for (;;) {
    double x = 0.0, result = 0.0;
    GoldenBracket next;
    if (!golden_probe(g, st, x, result, next)) return result;
    const double mx = value_at(x);
    const double mb = g.metric_b == kMax ? value_at(g.b) : g.metric_b;
    g = golden_update(next, x, mx, mb);
  }
}

// Metrics that can evaluate several transforms in one pass (values(ts)) and say whether that pays
// (batch_pays(): an extra transform in a pass costs less than a pass of its own).
template <typename M, typename = void>
struct has_batched_values : std::false_type {};
template <typename M>
struct has_batched_values<M, std::void_t<decltype(std::declval<M&>().values(std::declval<const std::vector<LinearMovingTransform>&>())),
                                         decltype(std::declval<const M&>().batch_pays())>> : std::true_type {};

// The same search, two probes per pass: with the probe x0 of the current bracket (and f(b) when it
// is not known yet) the pass also evaluates the probe that follows each outcome of x0 (f(x0) below
// f(b), or not). Which bracket follows depends on f(x0) only through that comparison, and the next
// probe on the bracket only, so the search visits exactly the brackets of golden_search, with the
// same values (a metric value is a function of the transform), and returns the same step; only the
// evaluations golden_search makes are counted. The probe of the outcome not taken is discarded.
template <typename Metric>
double golden_search_speculative(Metric& metric, const LinearMovingTransform& start, const std::vector<double>& dir,
                                 const CgSettings& st, double a, double b, double c, uint64_t& evaluations) {
  constexpr double kMax = std::numeric_limits<double>::max();
  const std::vector<double> base = start.optimised.parameters();
  auto at = [&](double step) {
    std::vector<double> p(base.size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = base[i] + step * dir[i];
    return start.with_parameters(p);
  };
  GoldenBracket g{a, b, c, kMax, 0};
  for (;;) {
    double x0 = 0.0, result = 0.0;
    GoldenBracket n0;
    if (!golden_probe(g, st, x0, result, n0)) return result;
    const bool need_b = g.metric_b == kMax;
    // the brackets after x0 if f(x0) < f(b) ("lower") and if not, f(x0) finite ("higher"): their
    // geometry does not depend on the values themselves
    const GoldenBracket lower = golden_update(n0, x0, 0.0, 1.0), higher = golden_update(n0, x0, 1.0, 0.0);
    double x1[2] = {0.0, 0.0}, r1[2] = {0.0, 0.0};
    GoldenBracket n1[2];
    const bool go[2] = {golden_probe(lower, st, x1[0], r1[0], n1[0]), golden_probe(higher, st, x1[1], r1[1], n1[1])};
    std::vector<LinearMovingTransform> ts{at(x0)};
    if (need_b) ts.push_back(at(g.b));
    int slot[2] = {-1, -1};
    for (int o = 0; o < 2; ++o)
      if (go[o]) {
        slot[o] = int(ts.size());
        ts.push_back(at(x1[o]));
      }
    const std::vector<double> v = metric.values(ts);
    const double mx = v[0];
    const double mb = need_b ? v[1] : g.metric_b;
    evaluations += need_b ? 2 : 1;
    const GoldenBracket g1 = golden_update(n0, x0, mx, mb);
    int o = -1;
    if (mx < mb) o = 0;
    else if (n0.c - n0.b > n0.b - n0.a || mx != kMax) o = 1;
    if (o < 0 || g1.metric_b == kMax) {
      // f(x0) or f(b) is the sentinel: the next step evaluates b again, which this pass did not
      g = g1;
      continue;
    }
    if (!go[o]) return r1[o];
    ++evaluations;
    g = golden_update(n1[o], x1[o], v[size_t(slot[o])], g1.metric_b);
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
    bool speculate = false;
    if constexpr (has_batched_values<Metric>::value) speculate = metric.batch_pays();
    if (speculate) {
      if constexpr (has_batched_values<Metric>::value)
        lr = golden_search_speculative(metric, current, conjugate, st, lr * st.lower_limit, lr, lr * st.upper_limit,
                                       trace.line_search_evaluations);
    } else {
      lr = golden_search(metric, current, conjugate, st, lr * st.lower_limit, lr, lr * st.upper_limit,
                         trace.line_search_evaluations);
    }
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
