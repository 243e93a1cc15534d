// Multi-resolution pyramid of antsRegistration (crates/fastwarp-reg/src/pyramid.rs): ANTs shrink
// factors per axis, ITK ShrinkImageFilter virtual grids, and the coefficients of ITK's recursive
// Gaussian (Deriche, fourth order, constant edge extension), whose recursions run on the device.
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

// Coefficients of ITK's zero-order recursive Gaussian (pyramid.rs:184-244, ITK
// RecursiveGaussianImageFilter): Deriche's fourth-order fit of the Gaussian (R. Deriche,
// "Recursively implementing the Gaussian and its derivatives", INRIA RR-1893, 1993),
//   g(t) ~ sum_r (a_r cos(w_r t) + b_r sin(w_r t)) exp(l_r t),  t >= 0 in units of sigma.
// Sampled every 1/sd (sd = sigma in voxels), each term is the impulse response of a two-pole
// section with numerator a + e (b sin - a cos) z^-1 and denominator 1 - 2 e cos z^-1 + e^2 z^-2
// (e = exp(l / sd), angles w / sd). The causal filter N/D is their sum over the common denominator
// D = D1 D2; the anti-causal filter is its mirror image without the k = 0 tap, so its numerator
// is M(z) = N(1/z) - n0 D(1/z) read as a polynomial in z. Both numerators are scaled so that the
// two filters together have unit DC gain. The lines are extended by their end values, so the
// outputs before a recursion starts are the steady state of that value: v N(1)/D(1) and
// v M(1)/D(1), which enter the first four steps as v * bn_k and v * bm_k.
struct Deriche {
  double n[4];   // causal numerator, taps x[i], x[i-1], x[i-2], x[i-3]
  double d[4];   // denominator without its leading 1, on y[i-1] .. y[i-4] (both directions)
  double m[4];   // anti-causal numerator, taps x[i+1] .. x[i+4]
  double bn[4];  // d_k times the causal steady-state gain
  double bm[4];  // d_k times the anti-causal steady-state gain

  static Deriche zero_order(double sigma, double spacing) {
    if (spacing < 1e-8) invalid("smoothing", "spacing is too small");
    const double sd = sigma / spacing;
    // Deriche's fitted constants (a, b, w, l) of the two terms, as ITK uses them
    static constexpr double kFit[2][4] = {{1.3530, 1.8151, 0.6681, -1.3932}, {-0.3531, 0.0902, 2.0787, -1.3732}};
    double num[2][2], den[2][3], num_dc[2], den_dc[2];
    for (int r = 0; r < 2; ++r) {
      const double a = kFit[r][0], b = kFit[r][1], w = kFit[r][2] / sd, e = std::exp(kFit[r][3] / sd);
      const double c = std::cos(w), s = std::sin(w);
      num[r][0] = a;
      num[r][1] = e * (b * s - a * c);
      den[r][0] = 1.0;
      den[r][1] = -2.0 * e * c;
      den[r][2] = e * e;
      // the sections at z = 1, without the cancellation in 1 - 2 e cos w + e^2 as the poles
      // approach 1 (large sigma): 1 - cos w = 2 sin^2(w / 2) and 1 - e = -expm1(l / sd)
      const double one_minus_e = -std::expm1(kFit[r][3] / sd);
      const double h = std::sin(0.5 * w);
      den_dc[r] = one_minus_e * one_minus_e + 4.0 * e * h * h;
      num_dc[r] = a * one_minus_e + e * (2.0 * a * h * h + b * s);
    }
    // D = D1 D2 and N = N1 D2 + N2 D1 as polynomials in z^-1
    double D[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, N[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) D[i + j] += den[0][i] * den[1][j];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) N[i + j] += num[0][i] * den[1][j] + num[1][i] * den[0][j];
    // N(1) / D(1): the causal filter's DC gain before scaling
    const double causal_gain = (num_dc[0] * den_dc[1] + num_dc[1] * den_dc[0]) / (den_dc[0] * den_dc[1]);
    // DC gain of causal plus anti-causal, the k = 0 tap counted once
    const double gain = 2.0 * causal_gain - N[0];
    Deriche k{};
    for (int q = 0; q < 4; ++q) {
      k.n[q] = N[q] / gain;
      k.d[q] = D[q + 1];
    }
    for (int q = 0; q < 4; ++q) k.m[q] = (q < 3 ? k.n[q + 1] : 0.0) - k.d[q] * k.n[0];
    // steady states of a constant unit input: N(1)/D(1) scaled, and M(1)/D(1) = that minus n0
    const double causal_dc = causal_gain / gain;
    const double anti_dc = causal_dc - k.n[0];
    for (int q = 0; q < 4; ++q) {
      k.bn[q] = k.d[q] * causal_dc;
      k.bm[q] = k.d[q] * anti_dc;
    }
    return k;
  }
};

}  // namespace fw
