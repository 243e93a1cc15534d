// ITK's WindowConvergenceMonitoringFunction (crates/fastwarp-reg/src/convergence.rs:47-126): the
// order-1, two-control-point B-spline fit of the last `window` normalised energies; the slope
// phi_0 - phi_1 is the convergence value.
#pragma once

#include <deque>

#include "common.hpp"

namespace fw {

class WindowConvergenceMonitor {
 public:
  explicit WindowConvergenceMonitor(size_t window) : window_(window) {}
  void add_energy_value(double v) {
    values_.push_back(v);
    if (values_.size() > window_) values_.pop_front();
    total_ += std::fabs(v);
  }
  void clear() {
    values_.clear();
    total_ = 0.0;
  }
  size_t size() const { return values_.size(); }

  // convergence.rs:90-125
  double convergence_value() const {
    if (values_.size() < window_) return std::numeric_limits<double>::max();
    if (window_ < 2) return 0.0;
    const double r = 1.0 / ((11.0 - 1.0) * 0.1);
    const double eps = r * 0.1 * 1e-3;
    double delta[2] = {0, 0}, omega[2] = {0, 0};
    size_t n = 0;
    for (double e : values_) {
      double p = (double(n) / double(window_ - 1)) * r;
      if (std::fabs(p - 1.0) <= eps) p = 1.0 - eps;
      if (p < 0.0 && std::fabs(p) <= eps) p = 0.0;
      const double t = p - std::floor(p);
      const double w[2] = {1.0 - t, t};
      const double w2 = w[0] * w[0] + w[1] * w[1];
      const double v = e / total_;
      for (int k = 0; k < 2; ++k) {
        omega[k] += w[k] * w[k];
        delta[k] += v * (w[k] * w[k] * w[k] / w2);
      }
      ++n;
    }
    double phi[2] = {0, 0};
    for (int k = 0; k < 2; ++k)
      if (std::fabs(omega[k]) > kEps) {
        const double q = delta[k] / omega[k];
        phi[k] = std::isfinite(q) ? q : 0.0;
      }
    return phi[0] - phi[1];
  }

 private:
  size_t window_;
  std::deque<double> values_;
  double total_ = 0.0;
};

}  // namespace fw
