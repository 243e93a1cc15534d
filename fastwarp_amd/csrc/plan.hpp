// Conversion between the C ABI's fw_plan / fw_result and the host recipes of one registration.
#pragma once

#include "../../include/fastwarp_hip.h"
#include "host/register.hpp"

namespace fw {

// The cohort recipes: Rigid/Affine `MI[...,1,32,Regular,0.25]`, `[1000x500x250x100,1e-6,10]`,
// `8x4x2x1`, `3x2x1x0vox` (linear.rs:112-145); `SyN[0.1,3,0]`, `CC[...,1,4]`,
// `[100x70x50x20,1e-6,10]`, `8x4x2x1`, `3x2x1x0vox` (recipe.rs:93-127).
inline void cohort_plan(fw_plan* p, int32_t channels) {
  *p = fw_plan{};
  const uint32_t lin_it[4] = {1000, 500, 250, 100}, syn_it[4] = {100, 70, 50, 20}, shrink[4] = {8, 4, 2, 1};
  const double sig[4] = {3.0, 2.0, 1.0, 0.0};
  p->linear_levels = p->syn_levels = 4;
  for (int l = 0; l < 4; ++l) {
    p->linear_iterations[l] = lin_it[l];
    p->linear_shrink[l] = p->syn_shrink[l] = shrink[l];
    p->linear_sigmas[l] = p->syn_sigmas[l] = sig[l];
    p->syn_iterations[l] = syn_it[l];
  }
  p->linear_gradient_step = 0.1;
  p->mi_bins = 32;
  p->sampling_percentage = 0.25;
  p->seed = 20260910u;
  p->linear_convergence_threshold = 1e-6;
  p->linear_convergence_window = 10;
  p->syn_gradient_step = 0.1;
  p->update_variance = 3.0;
  p->total_variance = 0.0;
  p->cc_radius = 4;
  p->channels = channels < 1 ? 1 : channels;
  for (int c = 0; c < p->channels && c < 8; ++c) p->channel_weights[c] = 1.0 / double(p->channels);
  p->syn_convergence_threshold = 1e-6;
  p->syn_convergence_window = 10;
  p->winsorize_lower = 0.005;
  p->winsorize_upper = 0.995;
  p->run_linear = 1;
  p->run_syn = 1;
}

inline RegisterPlan plan_from(const fw_plan* p) {
  if (!p) invalid("stage plan", "null plan");
  if (p->linear_levels < 1 || p->linear_levels > 8 || p->syn_levels < 1 || p->syn_levels > 8)
    invalid("stage plan", "between 1 and 8 levels per stage");
  if (p->channels < 1 || p->channels > 8) invalid("stage plan", "between 1 and 8 channels");
  RegisterPlan r;
  for (LinearRecipe* lr : {&r.rigid, &r.affine}) {
    lr->iterations.assign(p->linear_iterations, p->linear_iterations + p->linear_levels);
    lr->shrink_factors.assign(p->linear_shrink, p->linear_shrink + p->linear_levels);
    lr->smoothing_sigmas.assign(p->linear_sigmas, p->linear_sigmas + p->linear_levels);
    lr->gradient_step = p->linear_gradient_step;
    lr->bins = size_t(p->mi_bins);
    lr->sampling_percentage = float(p->sampling_percentage);
    lr->seed = p->seed;
    lr->convergence_threshold = p->linear_convergence_threshold;
    lr->convergence_window = size_t(p->linear_convergence_window);
  }
  r.rigid.kind = LinearKind::Rigid;
  r.affine.kind = LinearKind::Affine;
  SynRecipe& s = r.syn;
  s.iterations.assign(p->syn_iterations, p->syn_iterations + p->syn_levels);
  s.shrink_factors.assign(p->syn_shrink, p->syn_shrink + p->syn_levels);
  s.smoothing_sigmas.assign(p->syn_sigmas, p->syn_sigmas + p->syn_levels);
  s.gradient_step = p->syn_gradient_step;
  s.update_field_variance = p->update_variance;
  s.total_field_variance = p->total_variance;
  s.cc_radius = size_t(p->cc_radius);
  s.channel_weights.assign(p->channel_weights, p->channel_weights + p->channels);
  s.channel_metrics.clear();
  for (int32_t c = 0; c < p->channels; ++c) {
    const int32_t m = p->syn_channel_metrics[c];
    if (m != 0 && m != 1) invalid("stage plan", "a SyN channel metric is 0 (CC) or 1 (Mattes)");
    if (m == 1) s.channel_metrics.assign(size_t(p->channels), SynMetric::Cc);
  }
  if (!s.channel_metrics.empty())
    for (int32_t c = 0; c < p->channels; ++c)
      s.channel_metrics[c] = p->syn_channel_metrics[c] == 1 ? SynMetric::Mattes : SynMetric::Cc;
  s.mi_bins = size_t(p->mi_bins);
  s.convergence_threshold = p->syn_convergence_threshold;
  s.convergence_window = size_t(p->syn_convergence_window);
  r.winsorize_lower = float(p->winsorize_lower);
  r.winsorize_upper = float(p->winsorize_upper);
  r.run_linear = p->run_linear != 0;
  r.run_syn = p->run_syn != 0;
  // stages.rs:139-155: a given start must be finite and invertible
  switch (p->initial_kind) {
    case 0: r.initial_kind = InitialKind::CentreOfMass; break;
    case 1: r.initial_kind = InitialKind::Identity; break;
    case 2: {
      r.initial_kind = InitialKind::Affine;
      AffineTransform& t = r.initial_affine;
      bool finite = true;
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) t.matrix[i][j] = p->initial_affine.matrix[3 * i + j];
        t.translation[i] = p->initial_affine.translation[i];
        t.center[i] = p->initial_affine.center[i];
      }
      for (int i = 0; i < 3; ++i) {
        finite = finite && std::isfinite(t.translation[i]) && std::isfinite(t.center[i]);
        for (int j = 0; j < 3; ++j) finite = finite && std::isfinite(t.matrix[i][j]);
      }
      if (!finite) invalid("stage plan", "the initial transform has non-finite values");
      Mat3 inv;
      if (!mat3_inverse(t.matrix, inv)) invalid("stage plan", "the initial transform is singular");
      break;
    }
    default: invalid("stage plan", "initial_kind must be 0 (centre of mass), 1 (identity) or 2 (given affine)");
  }
  return r;
}

inline AffineTransform affine_of(const fw_affine* a) {
  AffineTransform t;
  if (!a) return t;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) t.matrix[i][j] = a->matrix[3 * i + j];
    t.translation[i] = a->translation[i];
    t.center[i] = a->center[i];
  }
  return t;
}

// One SyN stage (fw_syn_stage, and the oracle's counterpart) on the backend's images, taken as
// they are: the plan's SyN recipe from `initial`, with the stage's masks and earlier fields.
template <typename B>
RegisterOutput run_syn_stage(B& b, const fw_plan* plan, const std::vector<const typename B::ImageH*>& fixed,
                             const std::vector<const typename B::ImageH*>& moving, const fw_affine* initial,
                             const SynStageInputs<B>& extra, const volatile int32_t* cancel, FieldOutputs outputs = {},
                             std::optional<typename B::FieldH>* forward_device = nullptr,
                             std::optional<typename B::FieldH>* inverse_device = nullptr) {
  using ImageH = typename B::ImageH;
  if (!plan) invalid("syn stage", "null plan");
  const RegisterPlan p = plan_from(plan);
  if (fixed.size() != size_t(plan->channels) || moving.size() != size_t(plan->channels))
    invalid("syn", "the number of channels differs from the number of channel weights");
  std::vector<typename SmoothCache<B>::Pair> resident;
  for (size_t c = 0; c < fixed.size(); ++c) {
    if (!fixed[c] || !moving[c]) invalid("syn stage", "missing image");
    // the caller's images, not owned
    resident.push_back({std::shared_ptr<const ImageH>(fixed[c], [](const ImageH*) {}),
                        std::shared_ptr<const ImageH>(moving[c], [](const ImageH*) {})});
  }
  SmoothCache<B> cache(b, std::move(resident));
  RegisterOutput o;
  o.initial = o.total = affine_of(initial);
  const auto t = std::chrono::steady_clock::now();
  o.syn = syn_stage_of(b, p.syn, o.total, cache, cancel, outputs, forward_device, inverse_device, extra);
  o.seconds_syn = o.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
  return o;
}

// The per-iteration records of the calling thread's last registration (fw_last_trace).
inline std::vector<fw_trace_step>& last_trace() {
  static thread_local std::vector<fw_trace_step> t;
  return t;
}

inline void trace_to(const RegisterOutput& o, std::vector<fw_trace_step>& t) {
  t.clear();
  const LinearStageResult* st[2] = {&o.linear.rigid, &o.linear.affine};
  for (int s = 0; s < 2; ++s)
    for (size_t l = 0; l < st[s]->levels.size(); ++l) {
      const auto& values = st[s]->levels[l].cg.values;
      for (size_t i = 0; i < values.size(); ++i) {
        fw_trace_step r{};
        r.stage = s;
        r.level = int32_t(l);
        r.iteration = int32_t(i);
        r.value = values[i];
        t.push_back(r);
      }
    }
  for (size_t l = 0; l < o.syn.levels.size(); ++l)
    for (size_t i = 0; i < o.syn.levels[l].iterations.size(); ++i) {
      const SynIteration& it = o.syn.levels[l].iterations[i];
      fw_trace_step r{};
      r.stage = 2;
      r.level = int32_t(l);
      r.iteration = int32_t(i);
      r.value = it.energy;
      r.convergence_value = it.convergence_value;
      r.update_scale[0] = it.fixed_update_scale;
      r.update_scale[1] = it.moving_update_scale;
      for (int q = 0; q < 4; ++q) {
        r.inversion_steps[q] = int32_t(it.inversions[q].iterations);
        r.inversion_mean_error[q] = it.inversions[q].mean_error_norm;
        r.inversion_max_error[q] = it.inversions[q].max_error_norm;
      }
      t.push_back(r);
    }
}

inline int32_t copy_last_trace(fw_trace_step* out, int32_t capacity) {
  const auto& t = last_trace();
  const size_t n = out && capacity > 0 ? std::min(t.size(), size_t(capacity)) : 0;
  for (size_t i = 0; i < n; ++i) out[i] = t[i];
  return int32_t(t.size());
}

inline void result_to(const RegisterOutput& o, fw_result* r) {
  trace_to(o, last_trace());
  if (!r) return;
  *r = fw_result{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r->affine.matrix[3 * i + j] = o.total.matrix[i][j];
    r->affine.translation[i] = o.total.translation[i];
    r->affine.center[i] = o.total.center[i];
  }
  r->seconds_total = o.seconds_total;
  r->seconds_preprocess = o.seconds_preprocess;
  r->seconds_linear = o.seconds_linear;
  r->seconds_syn = o.seconds_syn;
  r->syn_levels = int32_t(std::min<size_t>(o.syn.levels.size(), 8));
  for (int l = 0; l < r->syn_levels; ++l) {
    r->syn_iterations[l] = int32_t(o.syn.levels[l].iterations.size());
    r->syn_last_energy[l] = o.syn.levels[l].iterations.empty() ? 0.0 : o.syn.levels[l].iterations.back().energy;
    r->syn_level_seconds[l] = o.syn.levels[l].seconds;
    double it = 0.0;
    int64_t steps = 0;
    for (const auto& i : o.syn.levels[l].iterations) {
      it += i.seconds;
      for (const InvertStats& st : i.inversions) steps += st.iterations;
    }
    r->syn_iteration_seconds[l] = it;
    r->syn_inversion_steps[l] = steps;
  }
  const LinearStageResult* st[2] = {&o.linear.rigid, &o.linear.affine};
  for (int s = 0; s < 2; ++s) {
    for (size_t l = 0; l < st[s]->levels.size() && l < 8; ++l) {
      const auto& lv = st[s]->levels[l];
      r->linear_iterations[s][l] = int32_t(lv.cg.iterations);
      r->linear_level_seconds[s][l] = lv.seconds;
      r->linear_iteration_seconds[s][l] = lv.cg.iteration_seconds;
      r->linear_evaluations[s][l] = int64_t(lv.cg.iterations + lv.cg.line_search_evaluations + (lv.cg.converged ? 1 : 0));
    }
    r->linear_final_value[s] = st[s]->levels.empty() || st[s]->levels.back().cg.values.empty() ? 0.0 : st[s]->levels.back().cg.values.back();
  }
  for (int i = 0; i < 6; ++i) r->rigid_parameters[i] = o.linear.rigid.transform.params[i];
  for (int i = 0; i < 12; ++i) r->affine_parameters[i] = o.linear.affine.transform.params[i];
}

inline Grid grid_of(const fw_grid* g) {
  if (!g) invalid("grid", "null grid");
  Mat3 dir;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dir[r][c] = g->direction[3 * r + c];
  for (int a = 0; a < 3; ++a)
    if (g->size[a] <= 0) invalid("grid", "zero size");
  return Grid({size_t(g->size[0]), size_t(g->size[1]), size_t(g->size[2])}, {g->spacing[0], g->spacing[1], g->spacing[2]},
              {g->origin[0], g->origin[1], g->origin[2]}, dir);
}

// Runs one registration and writes the fields (3n interleaved on the fixed grid) when asked.
template <typename B, typename E>
void run_register(B& backend, E& engine, const fw_plan* plan, const fw_grid* fixed_grid, const double* const* fixed,
                  const fw_grid* moving_grid, const double* const* moving, double* forward, double* inverse, fw_result* result,
                  const volatile int32_t* cancel) {
  const RegisterPlan p = plan_from(plan);
  const Grid fg = grid_of(fixed_grid), mg = grid_of(moving_grid);
  // the caller's arrays are read in place
  std::vector<ImageView> f, m;
  for (int32_t c = 0; c < plan->channels; ++c) {
    if (!fixed || !moving || !fixed[c] || !moving[c]) invalid("register", "null image data");
    f.emplace_back(fg, fixed[c]);
    m.emplace_back(mg, moving[c]);
  }
  const RegisterOutput o = register_full(backend, engine, p, f, m, cancel, FieldOutputs{forward, inverse});
  result_to(o, result);
}

}  // namespace fw
