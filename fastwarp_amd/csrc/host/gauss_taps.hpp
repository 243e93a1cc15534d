// ITK GaussianOperator coefficients (crates/fastwarp-reg/src/field_ops.rs:395-506): the taps the
// device smoothing kernels apply. Computed once per (variance, axis length) on the host in f64.
#pragma once

#include <cmath>
#include <vector>

namespace fw {

// ITK GaussianOperator::ModifiedBesselI0 (Abramowitz and Stegun polynomial), field_ops.rs:395-417.
inline double itk_bessel_i0(double y) {
  const double d = std::fabs(y);
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1 + m * 0.45813e-2)))));
  }
  const double m = 3.75 / d;
  return (std::exp(d) / std::sqrt(d)) *
         (0.39894228 + m * (0.1328592e-1 + m * (0.225319e-2 + m * (-0.157565e-2 + m * (0.916281e-2 + m * (-0.2057706e-1 +
                                                                                                      m * (0.2635537e-1 + m * (-0.1647633e-1 + m * 0.392377e-2))))))));
}

// ITK GaussianOperator::ModifiedBesselI1, field_ops.rs:420-440.
inline double itk_bessel_i1(double y) {
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

// ITK GaussianOperator::ModifiedBesselI for n >= 2 (Miller's downward recurrence normalised by
// I0), field_ops.rs:444-475.
inline double itk_bessel_in(int n, double y) {
  if (y == 0.0) return 0.0;
  const double toy = 2.0 / std::fabs(y);
  double qip = 0.0, acc = 0.0, qi = 1.0;
  for (int j = 2 * (n + int(std::sqrt(40.0 * double(n)))); j > 0; --j) {
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
  acc *= itk_bessel_i0(y) / qi;
  return (y < 0.0 && (n & 1)) ? -acc : acc;
}

// The 2r+1 symmetric taps of ITK's GaussianOperator for `variance` in squared voxels: the
// one-sided series exp(-t) I_n(t) until it reaches 1 - max_error or exceeds max_width, normalised
// (field_ops.rs:481-506).
inline std::vector<double> gaussian_operator_taps(double variance, double max_error, size_t max_width) {
  const double et = std::exp(-variance);
  std::vector<double> half{et * itk_bessel_i0(variance), et * itk_bessel_i1(variance)};
  double sum = half[0] + 2.0 * half[1];
  for (size_t i = 2; sum < 1.0 - max_error; ++i) {
    half.push_back(et * itk_bessel_in(int(i), variance));
    sum += 2.0 * half[i];
    if (half[i] <= 0.0 || half.size() > max_width) break;
  }
  std::vector<double> full;
  full.reserve(2 * half.size() - 1);
  for (size_t i = half.size() - 1; i >= 1; --i) full.push_back(half[i] / sum);
  for (double h : half) full.push_back(h / sum);
  return full;
}

}  // namespace fw
