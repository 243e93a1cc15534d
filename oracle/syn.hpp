// ORACLE — test infrastructure only. The CPU engine's own SyN stage: crates/fastwarp-reg/src/syn.rs
// restated directly over the oracle's f64 ops, independent of the product's SyN driver
// (fastwarp_amd/csrc/host/syn.hpp, which defers read-backs and pairs launches for the device). The
// whole-registration comparisons therefore check the product's driver as well as its ops. The
// registration pipeline picks this stage for the CPU engine by argument-dependent lookup of
// syn_stage_of.
#pragma once

#include <cmath>
#include <optional>
#include <vector>

#include "engine.hpp"

namespace fw::oracle {

namespace syn_detail {

// syn.rs:203-276 `SideAccumulator`: the values of one side, the weighted sum of the unit
// derivatives (or the single channel's derivative) and the sum of their magnitudes.
struct Side {
  std::vector<double> values;
  std::optional<VectorField> gradient;
  double total_magnitude = 0.0;

  void add(double weight, double value, VectorField derivative, std::optional<double> magnitude) {
    values.push_back(value);
    if (!magnitude) {
      gradient = std::move(derivative);
      return;
    }
    if (!gradient) gradient = VectorField::zeros(derivative.grid);
    if (!std::isfinite(*magnitude)) invalid("syn", "the metric gradient is not finite");
    total_magnitude += *magnitude;
    if (*magnitude > 2.220446049250313e-16) {
      const double a = weight / *magnitude;
      for (size_t i = 0; i < gradient->data.size(); ++i)
        for (int c = 0; c < 3; ++c) gradient->data[i][c] += a * derivative.data[i][c];
    }
  }
};

// syn.rs:248-270 `smoothed` and 285-300 `finish`, less the scaling (done by the caller once both
// maxima are known): the channel scaling, the first and weighted values, the update smoothing.
struct Smoothed {
  VectorField gradient;
  double first_value = 0.0, weighted_value = 0.0;
};

inline Smoothed smoothed(Side side, const SynRecipe& recipe, const std::vector<double>& normalised) {
  if (!side.gradient) invalid("syn", "no channel produced a gradient");
  Smoothed out;
  const size_t channels = side.values.size();
  out.first_value = side.values[0];
  VectorField g = std::move(*side.gradient);
  if (channels == 1) {
    out.weighted_value = out.first_value;
  } else {
    const double s = side.total_magnitude / double(channels);
    for (auto& v : g.data)
      for (int c = 0; c < 3; ++c) v[c] *= s;
    for (size_t c = 0; c < channels; ++c) out.weighted_value += side.values[c] * normalised[c];
  }
  out.gradient = smooth_field(g, recipe.update_field_variance);
  return out;
}

}  // namespace syn_detail

// syn.rs:776-986 `syn_fields` with the stage inputs of syn.rs:160-180 and 473-580.
inline SynFields oracle_syn_fields(CpuBackend& b, const SynRecipe& recipe, const AffineTransform& initial,
                                   SmoothCache<CpuBackend>& cache, const volatile int32_t* cancel, FieldOutputs outputs,
                                   std::optional<VectorField>* forward_device, std::optional<VectorField>* inverse_device,
                                   const SynStageInputs<CpuBackend>& extra) {
  using syn_detail::Side;
  recipe.validate();
  const size_t nc = cache.channels();
  if (nc == 0) invalid("syn", "no channels");
  if (nc != recipe.channel_weights.size())
    invalid("syn", "the number of channels differs from the number of channel weights");
  const Grid fixed_grid = cache.fixed_grid(0);
  for (size_t c = 0; c < nc; ++c)
    if (!cache.fixed_grid(c).approx_eq(fixed_grid, 1e-9)) invalid("syn", "fixed images of the channels differ in grid");
  if (bool(extra.prior) != bool(extra.prior_inverse))
    invalid("syn", "the earlier SyN stages need both their forward and their inverse field");
  if (extra.prior) {
    const Grid last = level_virtual_grid(
        fixed_grid, shrink_factors_per_dimension(recipe.shrink_factors[recipe.levels() - 1], fixed_grid.spacing));
    if (!(extra.prior->grid == last) || !(extra.prior_inverse->grid == last))
      invalid("syn",
              "the field of the earlier SyN stages is not on the grid of this stage's last level: every SyN stage must "
              "end with the same shrink factor");
  }
  double sum = 0.0;
  for (double w : recipe.channel_weights) sum += w;
  std::vector<double> normalised;
  for (double w : recipe.channel_weights) normalised.push_back(w / sum);
  bool cc = false, mattes = false;
  for (size_t c = 0; c < nc; ++c) (recipe.metric(c) == SynMetric::Cc ? cc : mattes) = true;
  const bool masks = extra.fixed_mask || extra.moving_mask;
  const VectorField* prior = extra.prior;
  const fw::InvertParams inv_params;  // InvertParams::SYN
  std::optional<VectorField> f2m, f2m_inv, m2m, m2m_inv;
  SynFields out;
  for (size_t level = 0; level < recipe.levels(); ++level) {
    check_cancel(cancel);
    const auto factors = shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing);
    const Grid grid = level_virtual_grid(fixed_grid, factors);
    const double sigma = recipe.smoothing_sigmas[level];
    std::vector<SmoothCache<CpuBackend>::Pair> smooth;
    for (size_t c = 0; c < nc; ++c) smooth.push_back(cache.get(c, sigma, recipe.sigma_units));
    if (!f2m) {
      f2m = VectorField::zeros(grid);
      f2m_inv = VectorField::zeros(grid);
      m2m = VectorField::zeros(grid);
      m2m_inv = VectorField::zeros(grid);
    } else {
      f2m = resample_field(*f2m, grid);
      f2m_inv = resample_field(*f2m_inv, grid);
      m2m = resample_field(*m2m, grid);
      m2m_inv = resample_field(*m2m_inv, grid);
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
      std::vector<Image> fmid, mmid;
      for (size_t c = 0; c < nc; ++c) {
        fmid.push_back(warp_through_field(*smooth[c].fixed, *f2m_inv, nullptr, nullptr));
        mmid.push_back(warp_through_field(*smooth[c].moving, *m2m_inv, prior, &initial));
      }
      // syn.rs:505-580: the masks on the level grid
      std::optional<Image> valid, fixed_in, moving_in;
      if (masks) {
        if (cc) {
          std::optional<Image> f;
          if (extra.fixed_mask) f = warp_mask_through_field(*extra.fixed_mask, *f2m_inv, nullptr, nullptr, nullptr);
          if (extra.moving_mask)
            valid = warp_mask_through_field(*extra.moving_mask, *m2m_inv, prior, &initial, f ? &*f : nullptr);
          else
            valid = std::move(f);
        }
        if (mattes) {
          if (extra.fixed_mask) fixed_in = warp_mask_through_field(*extra.fixed_mask, *f2m_inv, nullptr, nullptr, nullptr);
          if (extra.moving_mask) moving_in = warp_mask_through_field(*extra.moving_mask, *m2m_inv, prior, &initial, nullptr);
        }
      }
      // syn.rs:337-420 `side_updates`
      const bool single = nc == 1;
      Side fixed_side, moving_side;
      for (size_t c = 0; c < nc; ++c) {
        MetricPair<VectorField> pair =
            recipe.metric(c) == SynMetric::Cc
                ? b.cc_pair(fmid[c], mmid[c], valid ? &*valid : nullptr, recipe.cc_radius, !single)
                : b.mi_pair(fmid[c], mmid[c], fixed_in ? &*fixed_in : nullptr, moving_in ? &*moving_in : nullptr,
                            recipe.mi_bins, !single);
        fixed_side.add(normalised[c], pair.values[0], std::move(pair.fields[0]),
                       single ? std::nullopt : std::optional<double>(pair.norms[0]));
        moving_side.add(normalised[c], pair.values[1], std::move(pair.fields[1]),
                        single ? std::nullopt : std::optional<double>(pair.norms[1]));
      }
      syn_detail::Smoothed fs = syn_detail::smoothed(std::move(fixed_side), recipe, normalised);
      syn_detail::Smoothed ms = syn_detail::smoothed(std::move(moving_side), recipe, normalised);
      const double fmax = max_scaled_norm(fs.gradient), mmax = max_scaled_norm(ms.gradient);
      if (!(std::isfinite(fmax) && std::isfinite(mmax))) invalid("syn", "the metric gradient is not finite");
      SynIteration it;
      double scales[2] = {recipe.gradient_step, recipe.gradient_step};
      if (fmax > 0.0) scales[0] /= fmax;
      if (mmax > 0.0) scales[1] /= mmax;
      // syn.rs:422-445 `update_total_field`, the fixed side then the moving side
      syn_detail::Smoothed* sides[2] = {&fs, &ms};
      std::optional<VectorField>* fields[2] = {&f2m, &m2m};
      std::optional<VectorField>* inverses[2] = {&f2m_inv, &m2m_inv};
      for (int s = 0; s < 2; ++s) {
        VectorField& u = sides[s]->gradient;
        for (auto& v : u.data)
          for (int c = 0; c < 3; ++c) v[c] *= scales[s];
        VectorField tmp = smooth_field(compose(u, **fields[s]), recipe.total_field_variance);
        auto inv = b.invert(tmp, std::move(**inverses[s]), inv_params);
        auto fwd = b.invert(inv.first, std::move(tmp), inv_params);
        for (const fw::InvertStats* st : {&inv.second, &fwd.second})
          if (!(std::isfinite(st->mean_error_norm) && std::isfinite(st->max_error_norm)))
            invalid("syn", "the displacement field is not finite");
        *inverses[s] = std::move(inv.first);
        *fields[s] = std::move(fwd.first);
        it.inversions[2 * s] = inv.second;
        it.inversions[2 * s + 1] = fwd.second;
      }
      // syn.rs:941-958: the energy of the first metric into the monitor
      it.energy = 0.5 * (fs.first_value + ms.first_value);
      it.energy_weighted = 0.5 * (fs.weighted_value + ms.weighted_value);
      monitor.add_energy_value(it.energy);
      it.convergence_value = monitor.convergence_value();
      if (it.convergence_value < recipe.convergence_threshold) trace.converged = true;
      it.fixed_update_scale = scales[0];
      it.moving_update_scale = scales[1];
      trace.iterations.push_back(it);
    }
    out.levels.push_back(std::move(trace));
  }
  // syn.rs:969-986 and the collapse of syn.rs:746-774
  VectorField forward = compose(*m2m_inv, *f2m);
  VectorField inverse = compose(*f2m_inv, *m2m);
  if (prior) {
    forward = compose(*prior, forward);
    inverse = compose(inverse, *extra.prior_inverse);
  }
  if (outputs.forward) b.download_field_to(forward, outputs.forward);
  if (outputs.inverse) b.download_field_to(inverse, outputs.inverse);
  if (forward_device) *forward_device = std::move(forward);
  if (inverse_device) *inverse_device = std::move(inverse);
  return out;
}

// Whether the CPU engine runs the product's generic SyN loop over its ops instead of its own
// (a test compares the two).
inline bool& use_generic_syn() {
  static bool on = false;
  return on;
}

// The SyN stage of the CPU engine's registrations (found by argument-dependent lookup from the
// registration pipeline in place of the generic loop).
inline SynFields syn_stage_of(CpuBackend& b, const SynRecipe& recipe, const AffineTransform& initial,
                              SmoothCache<CpuBackend>& cache, const volatile int32_t* cancel, FieldOutputs outputs,
                              std::optional<VectorField>* forward_device, std::optional<VectorField>* inverse_device,
                              const SynStageInputs<CpuBackend>& extra) {
  if (use_generic_syn())
    return fw::syn_fields(b, recipe, initial, cache, cancel, outputs, forward_device, inverse_device, extra);
  return oracle_syn_fields(b, recipe, initial, cache, cancel, outputs, forward_device, inverse_device, extra);
}

}  // namespace fw::oracle
