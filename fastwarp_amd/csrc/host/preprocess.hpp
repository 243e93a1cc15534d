// antsRegistration intensity preprocessing (crates/fastwarp-reg/src/preprocess.rs): winsorise by
// the quantiles of a 256-bin histogram with ITK's marginal-scale margin, then window to [0, 1].
#pragma once

#include "grid.hpp"

namespace fw {

// preprocess.rs:40-150: the ImageToHistogramFilter histogram and its quantiles.
class IntensityHistogram {
 public:
  template <typename Values>
  explicit IntensityHistogram(const Values& values) {
    // range and counts over fixed chunks, merged in chunk order
    constexpr size_t kGrain = size_t(1) << 16;
    const size_t chunks = (values.size() + kGrain - 1) / kGrain;
    struct Range {
      double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
      size_t finite = 0;
    };
    std::vector<Range> ranges(chunks);
    parallel_chunks(values.size(), kGrain, [&](size_t c, size_t b, size_t e) {
      Range r;
      for (size_t i = b; i < e; ++i) {
        const double v = values[i];
        if (!std::isfinite(v)) continue;
        ++r.finite;
        r.lo = std::min(r.lo, v);
        r.hi = std::max(r.hi, v);
      }
      ranges[c] = r;
    });
    double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
    size_t finite = 0;
    for (const Range& r : ranges) {
      finite += r.finite;
      lo = std::min(lo, r.lo);
      hi = std::max(hi, r.hi);
    }
    if (finite == 0) invalid("preprocess", "image has no finite values");
    set_edges(lo, hi);
    std::vector<uint64_t> partial(chunks * kBins, 0);
    parallel_chunks(values.size(), kGrain, [&](size_t c, size_t b, size_t e) {
      uint64_t* counts = partial.data() + c * kBins;
      for (size_t i = b; i < e; ++i)
        if (std::isfinite(values[i])) ++counts[index_of(values[i])];
    });
    std::vector<uint64_t> counts(kBins, 0);
    for (size_t c = 0; c < chunks; ++c)
      for (size_t j = 0; j < kBins; ++j) counts[j] += partial[c * kBins + j];
    freq_.assign(counts.begin(), counts.end());
  }

  // From the finite range and the bin counts computed elsewhere (the device's histogram of the
  // same bins, see f32_thresholds).
  IntensityHistogram(double lo, double hi, const std::vector<uint64_t>& counts) {
    if (counts.size() != kBins) invalid("preprocess", "the histogram needs 256 counts");
    set_edges(lo, hi);
    freq_.assign(counts.begin(), counts.end());
  }

  // thresholds[k - 1] (k = 1..255) is the smallest f32 whose f64 value is at least the start of
  // bin k: for an f32 intensity v, index_of(v) is the number of thresholds at or below v.
  std::vector<float> f32_thresholds() const {
    std::vector<float> t(kBins - 1);
    for (size_t k = 1; k < kBins; ++k) {
      float f = float(min_[k]);
      if (double(f) < min_[k]) f = std::nextafter(f, std::numeric_limits<float>::infinity());
      t[k - 1] = f;
    }
    return t;
  }

  size_t index_of(double v) const {
    if (v < min_[0]) return 0;
    const size_t n = min_.size();
    if (v >= max_[n - 1]) return n - 1;
    // the largest j with min_[j] <= v (upper_bound - 1): start from the uniform-bin guess and walk
    const double guess = (v - min_[0]) / step_;
    size_t j = guess <= 0.0 ? 0 : std::min(n - 1, size_t(guess));
    while (j + 1 < n && v >= min_[j + 1]) ++j;
    while (j > 0 && v < min_[j]) --j;
    return j;
  }

  const std::vector<double>& bin_minima() const { return min_; }
  const std::vector<double>& bin_maxima() const { return max_; }

  double quantile(double p) const {
    const size_t size = freq_.size();
    double total = 0.0;
    for (double f : freq_) total += f;
    double cumulated = 0.0;
    if (p < 0.5) {
      size_t n = 0;
      double p_n = 0.0, p_prev = 0.0, f_n = 0.0;
      for (;;) {
        f_n = freq_[n];
        cumulated += f_n;
        p_prev = p_n;
        p_n = cumulated / total;
        ++n;
        if (!(n < size && p_n < p)) break;
      }
      const double proportion = f_n / total;
      const double mn = min_[n - 1];
      return mn + ((p - p_prev) / proportion) * (max_[n - 1] - mn);
    }
    long long n = (long long)size - 1;
    size_t m = 0;
    double p_n = 1.0, p_prev = 1.0, f_n = 0.0;
    for (;;) {
      f_n = freq_[size_t(n)];
      cumulated += f_n;
      p_prev = p_n;
      p_n = 1.0 - cumulated / total;
      --n;
      ++m;
      if (!(m < size && p_n > p)) break;
    }
    const double proportion = f_n / total;
    const size_t bin = size_t(n + 1);
    const double mx = max_[bin];
    return mx - ((p_prev - p) / proportion) * (mx - min_[bin]);
  }

 private:
  static constexpr size_t kBins = 256;
  // preprocess.rs:40-150: bin edges of [lo, hi + margin] in f32 steps, as ITK's histogram
  void set_edges(double lo, double hi) {
    const double margin = ((hi - lo) / double(kBins)) / 10.0;
    const double upper = hi + margin;
    const float interval = (float(upper) - float(lo)) / float(kBins);
    min_.assign(kBins, 0.0);
    max_.assign(kBins, 0.0);
    for (size_t j = 0; j + 1 < kBins; ++j) {
      min_[j] = lo + double(float(j) * interval);
      max_[j] = lo + double((float(j) + 1.0f) * interval);
    }
    min_[kBins - 1] = lo + double((float(kBins) - 1.0f) * interval);
    max_[kBins - 1] = upper;
    step_ = double(interval) > 0.0 ? double(interval) : 1.0;
  }
  std::vector<double> min_, max_, freq_;
  double step_ = 1.0;
};

// preprocess.rs:159-232: the quantile window rounded to f32 and the linear map of [lo, hi] onto
// [out_min, out_max].
struct IntensityWindow {
  double lo = 0.0, hi = 1.0, scale = 1.0, shift = 0.0, out_min = 0.0, out_max = 1.0;
};
inline IntensityWindow intensity_window(const IntensityHistogram& h, float lower, float upper, double out_min = 0.0,
                                        double out_max = 1.0) {
  IntensityWindow w;
  w.lo = double(float(h.quantile(double(lower))));
  w.hi = double(float(h.quantile(double(upper))));
  if (!(w.hi > w.lo)) invalid("preprocess", "degenerate intensity window");
  w.scale = (out_max - out_min) / (w.hi - w.lo);
  w.shift = out_min - w.lo * w.scale;
  w.out_min = out_min;
  w.out_max = out_max;
  return w;
}

// preprocess.rs:159-232: the quantile window rounded to f32, then the linear map of [lo, hi] onto
// [out_min, out_max] with clamping; non-finite voxels pass unchanged.
template <typename V>
Image preprocess_volume(const V& vol, float lower, float upper, double out_min = 0.0, double out_max = 1.0) {
  HostTimer host_timer("preprocess");
  IntensityHistogram h(vol.data);
  const IntensityWindow w = intensity_window(h, lower, upper, out_min, out_max);
  const double lo = w.lo, hi = w.hi, scale = w.scale, shift = w.shift;
  HostVec<double> out(vol.data.size());
  parallel_for(out.size(), [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const double x = vol.data[i];
      out[i] = x < lo ? out_min : (x > hi ? out_max : x * scale + shift);
    }
  });
  return Image(vol.grid, std::move(out));
}

}  // namespace fw
