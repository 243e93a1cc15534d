// ITK's symmetric normalisation as antsRegistration drives it (crates/fastwarp-reg/src/syn.rs),
// written once against a backend B with the stage interface of backend.rs:39-305:
//   B::ImageH / B::FieldH, upload_image, zero_field, resample_field, warp_through_field,
//   cc_pair, mi_pair, axpy, scale, smooth_field, max_scaled_norm_pair, compose, invert,
//   download_field, release_idle_buffers.
// Four fields live on the level grid (F, Finv, M, Minv); one iteration warps every channel's
// smoothed full-resolution images into the middle, evaluates the metric both ways, smooths and
// scales the two updates, composes them into the fields and re-inverts both fields.
#pragma once

#include <chrono>
#include <optional>

#include "backend_types.hpp"
#include "convergence.hpp"
#include "linear.hpp"
#include "pyramid.hpp"

namespace fw {

enum class SynMetric { Cc, Mattes };

// recipe.rs:52-127 `SynRecipe` (cohort: SyN[0.1,3,0], CC radius 4, 100x70x50x20, 8x4x2x1, 3x2x1x0vox).
struct SynRecipe {
  double gradient_step = 0.1;
  double update_field_variance = 3.0;
  double total_field_variance = 0.0;
  size_t cc_radius = 4;
  size_t mi_bins = 32;
  std::vector<SynMetric> channel_metrics;  // empty: CC for every channel
  std::vector<double> channel_weights{1.0};
  std::vector<uint32_t> iterations{100, 70, 50, 20};
  double convergence_threshold = 1e-6;
  size_t convergence_window = 10;
  std::vector<uint32_t> shrink_factors{8, 4, 2, 1};
  std::vector<double> smoothing_sigmas{3.0, 2.0, 1.0, 0.0};
  SigmaUnits sigma_units = SigmaUnits::Voxels;

  size_t levels() const { return iterations.size(); }
  SynMetric metric(size_t c) const { return channel_metrics.empty() ? SynMetric::Cc : channel_metrics[c]; }
  void validate() const {
    const size_t n = iterations.size();
    if (n == 0 || shrink_factors.size() != n || smoothing_sigmas.size() != n)
      invalid("syn recipe", "iterations, shrink factors and smoothing sigmas must have one entry per level");
    if (convergence_window < 2) invalid("syn recipe", "convergence window below 2");
    if (cc_radius == 0) invalid("syn recipe", "CC radius must be positive");
    if (channel_weights.empty()) invalid("syn recipe", "no channel weights");
    if (!channel_metrics.empty() && channel_metrics.size() != channel_weights.size())
      invalid("syn recipe", "one metric per channel is required");
  }
};

struct SynChannel {
  Image fixed, moving;
};

// syn.rs:96-128
struct SynIteration {
  double energy = 0.0, energy_weighted = 0.0, convergence_value = 0.0;
  double fixed_update_scale = 0.0, moving_update_scale = 0.0;
  InvertStats inversions[4];
  double seconds = 0.0;
};

struct SynLevelTrace {
  size_t level = 0;
  std::array<uint32_t, 3> shrink_factors{};
  std::array<size_t, 3> grid_size{};
  std::vector<SynIteration> iterations;
  bool converged = false;
  double seconds = 0.0;
};

// Per-level traces of one SyN run. The fields themselves go straight to the caller's arrays.
struct SynFields {
  std::vector<SynLevelTrace> levels;
};

// Where the SyN fields go: 3n interleaved f64 values each on the fixed grid; null skips one.
struct FieldOutputs {
  double* forward = nullptr;
  double* inverse = nullptr;
};

// What a SyN stage may start from beyond its linear transform (syn.rs:160-180 `SynStart::prior`,
// syn.rs:473-580 `StageMasks`): binary masks of the fixed and of the moving images (each on its own
// grid, non-zero inside; null for none) and the fields of the SyN stages before it (forward and
// inverse, on the grid of this stage's last level), which enter every moving-side warp and are
// collapsed with this stage's fields at the end.
template <typename B>
struct SynStageInputs {
  const typename B::ImageH* fixed_mask = nullptr;
  const typename B::ImageH* moving_mask = nullptr;
  const typename B::FieldH* prior = nullptr;
  const typename B::FieldH* prior_inverse = nullptr;
};

namespace detail {

// Backends that run both sides of a step in single launches (warp_pair, compose_pair,
// smooth_field_pair); the others get the two single operations, in the same order.
template <typename B, typename = void>
struct has_pair_ops : std::false_type {};
template <typename B>
struct has_pair_ops<B, std::void_t<decltype(&B::compose_pair), decltype(&B::warp_pair), decltype(&B::smooth_field_pair)>>
    : std::true_type {};

template <typename B>
std::pair<typename B::ImageH, typename B::ImageH> warp_both(B& b, const typename B::ImageH& f, const typename B::FieldH& finv,
                                                            const typename B::ImageH& m, const typename B::FieldH& minv,
                                                            const AffineTransform& initial) {
  if constexpr (has_pair_ops<B>::value) return b.warp_pair(f, finv, nullptr, m, minv, &initial);
  else return {b.warp_through_field(f, finv, nullptr, nullptr), b.warp_through_field(m, minv, nullptr, &initial)};
}

template <typename B>
std::pair<typename B::FieldH, typename B::FieldH> compose_both(B& b, const typename B::FieldH& u0, const typename B::FieldH& w0,
                                                               const typename B::FieldH& u1, const typename B::FieldH& w1) {
  if constexpr (has_pair_ops<B>::value) return b.compose_pair(u0, w0, u1, w1);
  else return {b.compose(u0, w0), b.compose(u1, w1)};
}

// Backends that scale the two updates while composing them (step_factors computes step / max on
// the device without scaling, compose_pair_scaled applies the factors as it interpolates).
template <typename B, typename = void>
struct has_scaled_compose : std::false_type {};
template <typename B>
struct has_scaled_compose<B, std::void_t<decltype(&B::step_factors), decltype(&B::compose_pair_scaled)>> : std::true_type {};

// Backends that invert the two sides' fields in the same launches (invert_pair_deferred).
template <typename B, typename = void>
struct has_invert_pair : std::false_type {};
template <typename B>
struct has_invert_pair<B, std::void_t<decltype(&B::invert_pair_deferred)>> : std::true_type {};

template <typename B>
std::pair<typename B::FieldH, typename B::FieldH> smooth_both(B& b, typename B::FieldH f0, typename B::FieldH f1, double variance) {
  if constexpr (has_pair_ops<B>::value) return b.smooth_field_pair(std::move(f0), std::move(f1), variance);
  else return {b.smooth_field(std::move(f0), variance), b.smooth_field(std::move(f1), variance)};
}

// syn.rs:203-297 `SideAccumulator` / `SmoothedSide` for one side of the update.
template <typename B>
struct Side {
  std::vector<double> values;
  std::optional<typename B::FieldH> gradient;
  double total_magnitude = 0.0;

  void add(B& b, double weight, double value, typename B::FieldH derivative, const double* magnitude) {
    values.push_back(value);
    if (!magnitude) {
      gradient = std::move(derivative);
      return;
    }
    if (!gradient) gradient = b.zero_field(b.field_grid(derivative));
    if (!std::isfinite(*magnitude)) invalid("syn", "the metric gradient is not finite");
    total_magnitude += *magnitude;
    if (*magnitude > kEps) b.axpy(*gradient, weight / *magnitude, derivative);
  }
};

}  // namespace detail

// syn.rs:776-986 `syn_fields` from a linear start `initial` (maps fixed points to moving points).
// The channels' images come from `cache` (its channels' fixed grids must agree). The fields go to
// the host buffers of `outputs` and, when given, stay on the device in *forward_device and
// *inverse_device.
template <typename B>
SynFields syn_fields(B& b, const SynRecipe& recipe, const AffineTransform& initial, SmoothCache<B>& cache,
                     const volatile int32_t* cancel, FieldOutputs outputs = {},
                     std::optional<typename B::FieldH>* forward_device = nullptr,
                     std::optional<typename B::FieldH>* inverse_device = nullptr, const SynStageInputs<B>& extra = {}) {
  using FieldH = typename B::FieldH;
  using ImageH = typename B::ImageH;
  recipe.validate();
  if (cache.channels() == 0) invalid("syn", "no channels");
  if (cache.channels() != recipe.channel_weights.size())
    invalid("syn", "the number of channels differs from the number of channel weights");
  const Grid fixed_grid = cache.fixed_grid(0);
  for (size_t c = 0; c < cache.channels(); ++c)
    if (!cache.fixed_grid(c).approx_eq(fixed_grid, 1e-9)) invalid("syn", "fixed images of the channels differ in grid");
  double wsum = 0.0;
  for (double w : recipe.channel_weights) wsum += w;
  std::vector<double> normalised;
  for (double w : recipe.channel_weights) normalised.push_back(w / wsum);
  const size_t nc = cache.channels();
  const InvertParams inv_params;  // InvertParams::SYN
  const typename B::FieldH* prior = extra.prior;
  if (bool(extra.prior) != bool(extra.prior_inverse))
    invalid("syn", "the earlier SyN stages need both their forward and their inverse field");
  if (prior) {
    // syn.rs:804-817
    const Grid last = level_virtual_grid(
        fixed_grid, shrink_factors_per_dimension(recipe.shrink_factors[recipe.levels() - 1], fixed_grid.spacing));
    if (!(b.field_grid(*extra.prior) == last) || !(b.field_grid(*extra.prior_inverse) == last))
      invalid("syn",
              "the field of the earlier SyN stages is not on the grid of this stage's last level: every SyN stage must "
              "end with the same shrink factor");
  }
  const bool masks = extra.fixed_mask || extra.moving_mask;
  bool uses_cc = false, uses_mattes = false;
  for (size_t c = 0; c < nc; ++c) (recipe.metric(c) == SynMetric::Cc ? uses_cc : uses_mattes) = true;
  std::optional<FieldH> f2m, f2m_inv, m2m, m2m_inv;
  SynFields out;
  for (size_t level = 0; level < recipe.levels(); ++level) {
    check_cancel(cancel);
    const auto t0 = std::chrono::steady_clock::now();
    const auto factors = shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing);
    const Grid grid = level_virtual_grid(fixed_grid, factors);
    const double sigma = recipe.smoothing_sigmas[level];
    std::vector<typename SmoothCache<B>::Pair> smooth;
    for (size_t c = 0; c < nc; ++c) smooth.push_back(cache.get(c, sigma, recipe.sigma_units));
    if (!f2m) {
      f2m = b.zero_field(grid);
      f2m_inv = b.zero_field(grid);
      m2m = b.zero_field(grid);
      m2m_inv = b.zero_field(grid);
    } else {
      // syn.rs:861-882: resample the four fields onto the next level grid
      f2m = b.resample_field(*f2m, grid);
      f2m_inv = b.resample_field(*f2m_inv, grid);
      m2m = b.resample_field(*m2m, grid);
      m2m_inv = b.resample_field(*m2m_inv, grid);
    }
    WindowConvergenceMonitor monitor(recipe.convergence_window);
    SynLevelTrace trace;
    trace.level = level;
    trace.shrink_factors = factors;
    trace.grid_size = grid.size;
    uint32_t current = 0;
    while (current < recipe.iterations[level] && !trace.converged) {
      check_cancel(cancel);
      ++current;
      const auto t_it = std::chrono::steady_clock::now();
      // The iteration is issued without waiting on the device: the metric values, the update
      // maxima and the inversion statistics are read back once, at the settle below, where the
      // reference's finiteness checks run on them (same errors, after the device work).
      // the masks of this iteration on the level grid (syn.rs:505-580 `resolve`): for CC the
      // voxels inside both, for Mattes each mask on its own
      std::optional<ImageH> valid, fixed_in, moving_in;
      if (masks) {
        if (uses_cc) {
          std::optional<ImageH> f;
          if (extra.fixed_mask) f = b.warp_mask_through_field(*extra.fixed_mask, *f2m_inv, nullptr, nullptr, nullptr);
          if (extra.moving_mask)
            valid = b.warp_mask_through_field(*extra.moving_mask, *m2m_inv, prior, &initial, f ? &*f : nullptr);
          else
            valid = std::move(f);
        }
        if (uses_mattes) {
          if (extra.fixed_mask) fixed_in = b.warp_mask_through_field(*extra.fixed_mask, *f2m_inv, nullptr, nullptr, nullptr);
          if (extra.moving_mask)
            moving_in = b.warp_mask_through_field(*extra.moving_mask, *m2m_inv, prior, &initial, nullptr);
        }
      }
      // 1. resample every channel into the middle (syn.rs:895-911)
      detail::Side<B> fs, ms;
      const bool single = nc == 1;
      std::vector<Pending<std::array<double, 2>>> later(nc);
      for (size_t c = 0; c < nc; ++c) {
        std::pair<ImageH, ImageH> mids =
            prior ? std::pair<ImageH, ImageH>{b.warp_through_field(*smooth[c].fixed, *f2m_inv, nullptr, nullptr),
                                              b.warp_through_field(*smooth[c].moving, *m2m_inv, prior, &initial)}
                  : detail::warp_both(b, *smooth[c].fixed, *f2m_inv, *smooth[c].moving, *m2m_inv, initial);
        ImageH fmid = std::move(mids.first), mmid = std::move(mids.second);
        // 2. the metric both ways (syn.rs:337-420 `side_updates`); several channels need the
        // derivative norms for their weights at once, a single one needs its values only later
        MetricPair<FieldH> pair;
        if (recipe.metric(c) == SynMetric::Cc && single && !valid) {
          pair = b.cc_pair_deferred(fmid, mmid, recipe.cc_radius);
          later[c] = std::move(pair.later);
        } else {
          pair = recipe.metric(c) == SynMetric::Cc
                     ? b.cc_pair(fmid, mmid, valid ? &*valid : nullptr, recipe.cc_radius, !single)
                     : b.mi_pair(fmid, mmid, fixed_in ? &*fixed_in : nullptr, moving_in ? &*moving_in : nullptr,
                                 recipe.mi_bins, !single);
          later[c] = Pending<std::array<double, 2>>(std::array<double, 2>{pair.values[0], pair.values[1]});
        }
        fs.add(b, normalised[c], 0.0, std::move(pair.fields[0]), single ? nullptr : &pair.norms[0]);
        ms.add(b, normalised[c], 0.0, std::move(pair.fields[1]), single ? nullptr : &pair.norms[1]);
      }
      FieldH grads[2];
      detail::Side<B>* sides[2] = {&fs, &ms};
      for (int s = 0; s < 2; ++s) {
        detail::Side<B>& side = *sides[s];
        if (!side.gradient) invalid("syn", "no channel produced a gradient");
        if (nc > 1) b.scale(*side.gradient, side.total_magnitude / double(nc));
      }
      {
        auto smoothed = detail::smooth_both(b, std::move(*fs.gradient), std::move(*ms.gradient), recipe.update_field_variance);
        grads[0] = std::move(smoothed.first);
        grads[1] = std::move(smoothed.second);
      }
      // 3. scale each update so its largest voxel step is the gradient step, and 4-5. compose
      // into the total fields and re-estimate both inverses (syn.rs:422-445)
      SynIteration it;
      std::optional<FieldH>* fields[2] = {&f2m, &m2m};
      std::optional<FieldH>* inverses[2] = {&f2m_inv, &m2m_inv};
      Pending<InvertStats> stats[4];
      Pending<std::pair<double, double>> maxima;
      std::pair<FieldH, FieldH> composed;
      if constexpr (detail::has_scaled_compose<B>::value) {
        auto factors = b.step_factors(grads[0], grads[1], recipe.gradient_step);
        maxima = std::move(factors.first);
        composed = b.compose_pair_scaled(grads[0], **fields[0], grads[1], **fields[1], factors.second);
      } else {
        maxima = b.scale_to_step(grads[0], grads[1], recipe.gradient_step);
        composed = detail::compose_both(b, grads[0], **fields[0], grads[1], **fields[1]);
      }
      fields[0]->reset();
      fields[1]->reset();
      auto totals = detail::smooth_both(b, std::move(composed.first), std::move(composed.second), recipe.total_field_variance);
      FieldH tmps[2] = {std::move(totals.first), std::move(totals.second)};
      if constexpr (detail::has_invert_pair<B>::value) {
        // both sides at once: the new inverses from the old ones, then the new fields from them
        auto inv = b.invert_pair_deferred(tmps[0], std::move(**inverses[0]), tmps[1], std::move(**inverses[1]), inv_params);
        auto fwd = b.invert_pair_deferred(inv.first, std::move(tmps[0]), inv.second, std::move(tmps[1]), inv_params);
        *inverses[0] = std::move(inv.first);
        *inverses[1] = std::move(inv.second);
        *fields[0] = std::move(fwd.first);
        *fields[1] = std::move(fwd.second);
        stats[0] = std::move(inv.first_stats);
        stats[1] = std::move(fwd.first_stats);
        stats[2] = std::move(inv.second_stats);
        stats[3] = std::move(fwd.second_stats);
      } else {
        for (int s = 0; s < 2; ++s) {
          FieldH tmp = std::move(tmps[s]);
          auto inv = b.invert_deferred(tmp, std::move(**inverses[s]), inv_params);
          auto fwd = b.invert_deferred(inv.first, std::move(tmp), inv_params);
          *inverses[s] = std::move(inv.first);
          *fields[s] = std::move(fwd.first);
          stats[2 * s] = std::move(inv.second);
          stats[2 * s + 1] = std::move(fwd.second);
        }
      }
      b.settle();
      const auto [fmax, mmax] = maxima.get();
      if (!(std::isfinite(fmax) && std::isfinite(mmax))) invalid("syn", "the metric gradient is not finite");
      for (int q = 0; q < 4; ++q) {
        it.inversions[q] = stats[q].get();
        if (!(std::isfinite(it.inversions[q].mean_error_norm) && std::isfinite(it.inversions[q].max_error_norm)))
          invalid("syn", "the displacement field is not finite");
      }
      double weighted[2] = {0.0, 0.0};
      for (size_t c = 0; c < nc; ++c) {
        const std::array<double, 2> v = later[c].get();
        fs.values[c] = v[0];
        ms.values[c] = v[1];
        weighted[0] += single ? v[0] : v[0] * normalised[c];
        weighted[1] += single ? v[1] : v[1] * normalised[c];
      }
      double scales[2] = {recipe.gradient_step, recipe.gradient_step};
      if (fmax > 0.0) scales[0] /= fmax;
      if (mmax > 0.0) scales[1] /= mmax;
      // 6. energy of the first metric into the monitor (syn.rs:941-948)
      it.energy = 0.5 * (fs.values[0] + ms.values[0]);
      it.energy_weighted = 0.5 * (weighted[0] + weighted[1]);
      monitor.add_energy_value(it.energy);
      it.convergence_value = monitor.convergence_value();
      if (it.convergence_value < recipe.convergence_threshold) trace.converged = true;
      it.fixed_update_scale = scales[0];
      it.moving_update_scale = scales[1];
      it.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
      trace.iterations.push_back(it);
    }
    trace.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    out.levels.push_back(std::move(trace));
  }
  // syn.rs:969-980: SyN = Minv o F forward, Finv o M inverse; after earlier stages, collapsed with
  // theirs (syn.rs:746-774): forward earlier(x + later(x)) + later(x), inverse
  // later_inverse(x + earlier_inverse(x)) + earlier_inverse(x)
  if (outputs.forward || forward_device) {
    FieldH f = b.compose(*m2m_inv, *f2m);
    if (prior) f = b.compose(*prior, f);
    if (outputs.forward) b.download_field_to(f, outputs.forward);
    if (forward_device) *forward_device = std::move(f);
  }
  if (outputs.inverse || inverse_device) {
    FieldH f = b.compose(*f2m_inv, *m2m);
    if (prior) f = b.compose(f, *extra.prior_inverse);
    if (outputs.inverse) b.download_field_to(f, outputs.inverse);
    if (inverse_device) *inverse_device = std::move(f);
  }
  return out;
}

// The SyN stage a registration runs on backend B: the loop above. An engine may provide its own
// stage for its backend type (an overload found by argument-dependent lookup).
template <typename B>
SynFields syn_stage_of(B& b, const SynRecipe& recipe, const AffineTransform& initial, SmoothCache<B>& cache,
                       const volatile int32_t* cancel, FieldOutputs outputs,
                       std::optional<typename B::FieldH>* forward_device, std::optional<typename B::FieldH>* inverse_device,
                       const SynStageInputs<B>& extra) {
  return syn_fields(b, recipe, initial, cache, cancel, outputs, forward_device, inverse_device, extra);
}

}  // namespace fw
