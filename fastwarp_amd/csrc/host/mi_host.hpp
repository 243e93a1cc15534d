// Host half of Mattes mutual information: bin geometry and the f64 finish of an accumulated joint
// histogram (crates/fastwarp-reg/src/metric/mi.rs:121-317 `MiBinning`, :702-769
// `finish_histogram`). The device builds the integer histograms; the host turns them into the
// value and the per-cell derivative weights, as the reference's GPU engine does
// (crates/fastwarp-gpu/src/wgpu_backend/mi.rs:418-490).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "common.hpp"

namespace fw {

struct MiBinning {
  size_t bins = 32;
  double fixed_min = 0, fixed_max = 0, moving_min = 0, moving_max = 0;
  double fixed_bin_size = 0, moving_bin_size = 0, fixed_normalized_min = 0, moving_normalized_min = 0;

  // mi.rs:238-276 `from_ranges`
  static MiBinning from_ranges(size_t bins, double fmin, double fmax, double mmin, double mmax) {
    if (bins <= 4) invalid("mattes", "at least five histogram bins are required");
    if (fmax - fmin <= kEps) invalid("mattes", "the fixed image is constant");
    if (mmax - mmin <= kEps) invalid("mattes", "the moving image is constant");
    MiBinning b;
    b.bins = bins;
    b.fixed_min = fmin;
    b.fixed_max = fmax;
    b.moving_min = mmin;
    b.moving_max = mmax;
    const double span = double(bins - 4);
    b.fixed_bin_size = (fmax - fmin) / span;
    b.moving_bin_size = (mmax - mmin) / span;
    b.fixed_normalized_min = fmin / b.fixed_bin_size - 2.0;
    b.moving_normalized_min = mmin / b.moving_bin_size - 2.0;
    return b;
  }
  // mi.rs:307-316: the truncated term clamped to [2, bins - 3].
  size_t clamp_bin(double term) const {
    // saturating truncation, as Rust's `as i64`
    long long idx = std::isnan(term) ? 0 : (long long)std::max(-1e18, std::min(1e18, term));
    if (idx < 2) idx = 2;
    if (idx > (long long)bins - 3) idx = (long long)bins - 3;
    return size_t(idx);
  }
  size_t fixed_bin(double v) const { return clamp_bin(v / fixed_bin_size - fixed_normalized_min); }

  // The f32 thresholds of the fixed-axis bins: thresholds[b - 3] is the smallest f32 value whose
  // f64 fixed bin is at least b (b = 3 .. bins - 3), so that a kernel that counts the thresholds
  // at or below an f32 intensity reproduces the f64 bin exactly.
  std::vector<float> fixed_thresholds() const {
    std::vector<float> t;
    for (size_t b = 3; b + 3 <= bins; ++b) {
      // fixed_bin is monotone in v: bisect the ordered f32 values between the range ends
      auto key = [](float f) {
        int32_t i;
        std::memcpy(&i, &f, 4);
        return i < 0 ? int64_t(INT32_MIN) - int64_t(i) : int64_t(i);
      };
      auto from_key = [](int64_t k) {
        int32_t i = k < 0 ? int32_t(int64_t(INT32_MIN) - k) : int32_t(k);
        float f;
        std::memcpy(&f, &i, 4);
        return f;
      };
      int64_t lo = key(-std::numeric_limits<float>::max()), hi = key(std::numeric_limits<float>::max());
      if (fixed_bin(double(from_key(hi))) < b) {
        t.push_back(std::numeric_limits<float>::infinity());
        continue;
      }
      while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (fixed_bin(double(from_key(mid))) >= b) hi = mid;
        else lo = mid;
      }
      t.push_back(from_key(hi));
    }
    return t;
  }
};

struct HistogramFinish {
  double value = 0.0;
  std::vector<double> ratio;  // nFactor * log(p(f,m)/p_M(m)) per cell, zero where ITK skips it
};

// mi.rs:702-769: Kahan-normalised PDF, marginals, -MI and the derivative weights.
inline HistogramFinish finish_histogram(size_t bins, double moving_bin_size, uint64_t num_valid, std::vector<double> joint,
                                        std::vector<double> fixed_marginal, bool want_ratio) {
  if (num_valid == 0) invalid("mattes", "all samples map outside the moving image buffer");
  // Kahan sum. A zero term with sum - comp == sum in f64 leaves both unchanged (y = -comp,
  // t = sum, comp' = (t - sum) - y = comp), so such terms are skipped: most of a joint
  // histogram is empty, and the result is that of the full sequence.
  double sum = 0.0, comp = 0.0;
  for (double v : joint) {
    if (v == 0.0 && sum - comp == sum) continue;
    const double y = v - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  if (sum < kEps) invalid("mattes", "joint PDF summed to zero");
  const double norm = 1.0 / sum;
  for (double& v : joint) v *= norm;
  for (double& v : fixed_marginal) v *= norm;
  std::vector<double> mm(bins, 0.0);
  for (size_t r = 0; r < bins; ++r)
    for (size_t c = 0; c < bins; ++c) mm[c] += joint[r * bins + c];
  const double n_factor = -1.0 / (moving_bin_size * double(num_valid));
  HistogramFinish out;
  if (want_ratio) out.ratio.assign(bins * bins, 0.0);
  double total = 0.0;
  for (size_t f = 0; f < bins; ++f) {
    const double fm = fixed_marginal[f];
    if (fm <= kEps) continue;
    const double lf = std::log(fm);
    for (size_t m = 0; m < bins; ++m) {
      const double jv = joint[f * bins + m];
      if (mm[m] > kEps && jv > kEps) {
        const double pr = std::log(jv / mm[m]);
        total += jv * (pr - lf);
        if (want_ratio) out.ratio[f * bins + m] = n_factor * pr;
      }
    }
  }
  out.value = -total;
  return out;
}

}  // namespace fw
