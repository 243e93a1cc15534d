// extern "C" entry points of include/fastwarp_hip.h over HipBackend. Each function catches every
// exception and turns it into a status code plus the context's last-error text, as the trait
// methods return Result<_, RegError> (crates/fastwarp-reg/src/error.rs).
#include "../../include/fastwarp_hip.h"

#include <cstring>
#include <string>

#include "hip/backend.hpp"
#include "hip/batch.hpp"
#include "host/capacity.hpp"
#include "host/itk_h5.hpp"
#include "host/itk_mat.hpp"
#include "plan.hpp"

using fw::Grid;
using fw::hip::HField;
using fw::hip::HImage;
using fw::hip::HipBackend;

struct fw_ctx {
  HipBackend backend;
  fw::hip::WorkerPool workers;  // the side-by-side linear fits of the batched entry points
  std::string error;
  explicit fw_ctx(int device) : backend(device), workers(device) {}
};
struct fw_image {
  HImage h;
};
struct fw_field {
  HField h;
};
struct fw_mi_level {
  fw_ctx* ctx;
  fw::hip::MiLevel level;
};

namespace {

thread_local std::string g_create_error;

Grid grid_from(const fw_grid* g) {
  if (!g) fw::invalid("grid", "null grid");
  for (int a = 0; a < 3; ++a)
    if (g->size[a] <= 0) fw::invalid("grid", "zero size");
  fw::Mat3 dir;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dir[r][c] = g->direction[3 * r + c];
  return Grid({size_t(g->size[0]), size_t(g->size[1]), size_t(g->size[2])}, {g->spacing[0], g->spacing[1], g->spacing[2]},
              {g->origin[0], g->origin[1], g->origin[2]}, dir);
}

void grid_to(const Grid& g, fw_grid* out) {
  for (int a = 0; a < 3; ++a) {
    out->size[a] = int64_t(g.size[a]);
    out->spacing[a] = g.spacing[a];
    out->origin[a] = g.origin[a];
    for (int c = 0; c < 3; ++c) out->direction[3 * a + c] = g.direction[a][c];
  }
}

fw::AffineTransform affine_from(const fw_affine* a) {
  fw::AffineTransform t;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.matrix[r][c] = a->matrix[3 * r + c];
    t.translation[r] = a->translation[r];
    t.center[r] = a->center[r];
  }
  return t;
}

template <typename F>
int guard(fw_ctx* ctx, F&& f) {
  if (!ctx) return FW_ERR_INVALID;
  try {
    f();
    ctx->error.clear();
    return FW_OK;
  } catch (const fw::Cancelled& e) {
    ctx->error = e.what();
    return FW_ERR_CANCELLED;
  } catch (const fw::Error& e) {
    ctx->error = e.what();
    return e.what_ == "hip" ? FW_ERR_DEVICE : FW_ERR_INVALID;
  } catch (const std::exception& e) {
    ctx->error = e.what();
    return FW_ERR_INVALID;
  }
}

}  // namespace

namespace fw {
// The message fw_last_error(NULL) returns for a call with no context.
void set_call_error(const std::string& message) { g_create_error = message; }
}  // namespace fw

namespace {

void need(const void* p, const char* what) {
  if (!p) fw::invalid("fastwarp", std::string(what) + " is null");
}

}  // namespace

extern "C" {

int fw_ctx_create(int device, fw_ctx** out) {
  if (!out) return FW_ERR_INVALID;
  try {
    *out = new fw_ctx(device);
    return FW_OK;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    *out = nullptr;
    return FW_ERR_DEVICE;
  }
}

void fw_ctx_destroy(fw_ctx* ctx) { delete ctx; }

const char* fw_last_error(const fw_ctx* ctx) { return ctx ? ctx->error.c_str() : g_create_error.c_str(); }

int fw_release_idle(fw_ctx* ctx) {
  return guard(ctx, [&] { ctx->backend.release_idle_buffers(); });
}

int fw_set_tuning(fw_ctx* ctx, const char* key, int64_t value) {
  return guard(ctx, [&] {
    need(key, "key");
    const std::string k(key);
    const std::string knob = k.substr(0, k.find('@'));
    if (value >= 0 && !fw::TuningTable::valid(knob, value))
      fw::invalid("tuning", "unknown knob or value out of range: " + k + " = " + std::to_string(value));
    ctx->backend.ctx().tuning().set(k, value);
  });
}

int fw_memory_stats(fw_ctx* ctx, uint64_t* held, uint64_t* peak) {
  return guard(ctx, [&] {
    if (held) *held = ctx->backend.ctx().held();
    if (peak) *peak = ctx->backend.ctx().peak();
  });
}

int fw_image_upload(fw_ctx* ctx, const fw_grid* grid, const double* data, fw_image** out) {
  return guard(ctx, [&] {
    need(data, "image data");
    need(out, "output");
    const Grid g = grid_from(grid);
    fw::Image v(g, std::vector<double>(data, data + g.num_voxels()));
    *out = new fw_image{ctx->backend.upload_image(v)};
  });
}

int fw_image_upload_f32(fw_ctx* ctx, const fw_grid* grid, const float* data, fw_image** out) {
  return guard(ctx, [&] {
    need(data, "image data");
    need(out, "output");
    *out = new fw_image{ctx->backend.upload_image_f32(grid_from(grid), data)};
  });
}

int fw_image_download(fw_ctx* ctx, const fw_image* image, double* out) {
  return guard(ctx, [&] {
    need(image, "image");
    need(out, "output");
    fw::Image v = ctx->backend.download_image(image->h);
    std::memcpy(out, v.data.data(), v.data.size() * sizeof(double));
  });
}

int fw_resample_isotropic(fw_ctx* ctx, const fw_image* image, double mm, fw_image** out) {
  return guard(ctx, [&] {
    need(image, "image");
    need(out, "output");
    *out = new fw_image{ctx->backend.resample_isotropic(image->h, mm)};
  });
}

int fw_image_grid(const fw_image* image, fw_grid* out) {
  if (!image || !out) return FW_ERR_INVALID;
  grid_to(image->h.grid, out);
  return FW_OK;
}

void fw_image_free(fw_ctx*, fw_image* image) { delete image; }

int fw_field_zero(fw_ctx* ctx, const fw_grid* grid, fw_field** out) {
  return guard(ctx, [&] {
    need(out, "output");
    *out = new fw_field{ctx->backend.zero_field(grid_from(grid))};
  });
}

int fw_field_upload(fw_ctx* ctx, const fw_grid* grid, const double* aos, fw_field** out) {
  return guard(ctx, [&] {
    need(aos, "field data");
    need(out, "output");
    const Grid g = grid_from(grid);
    std::vector<fw::Vec3> d(g.num_voxels());
    for (size_t i = 0; i < d.size(); ++i) d[i] = {aos[3 * i], aos[3 * i + 1], aos[3 * i + 2]};
    *out = new fw_field{ctx->backend.upload_field(fw::VectorField(g, std::move(d)))};
  });
}

int fw_field_download(fw_ctx* ctx, const fw_field* field, double* aos) {
  return guard(ctx, [&] {
    need(field, "field");
    need(aos, "output");
    fw::VectorField v = ctx->backend.download_field(field->h);
    for (size_t i = 0; i < v.data.size(); ++i)
      for (int c = 0; c < 3; ++c) aos[3 * i + c] = v.data[i][c];
  });
}

int fw_field_grid(const fw_field* field, fw_grid* out) {
  if (!field || !out) return FW_ERR_INVALID;
  grid_to(field->h.grid, out);
  return FW_OK;
}

int fw_field_copy(fw_ctx* ctx, const fw_field* field, fw_field** out) {
  return guard(ctx, [&] {
    need(field, "field");
    need(out, "output");
    *out = new fw_field{ctx->backend.copy_field(field->h)};
  });
}

void fw_field_free(fw_ctx*, fw_field* field) { delete field; }

int fw_smooth_image(fw_ctx* ctx, const fw_image* src, const double sigma_mm[3], fw_image** out) {
  return guard(ctx, [&] {
    need(src, "source");
    need(sigma_mm, "sigma");
    need(out, "output");
    *out = new fw_image{ctx->backend.smooth_image_mm(src->h, {sigma_mm[0], sigma_mm[1], sigma_mm[2]})};
  });
}

int fw_preprocess_image(fw_ctx* ctx, const fw_image* raw, float lower, float upper, fw_image** out) {
  return guard(ctx, [&] {
    need(raw, "image");
    need(out, "output");
    *out = new fw_image{ctx->backend.preprocess_image(raw->h, lower, upper)};
  });
}

int fw_centre_of_gravity(fw_ctx* ctx, const fw_image* img, double centre[3]) {
  return guard(ctx, [&] {
    need(img, "image");
    need(centre, "output");
    const fw::Vec3 c = ctx->backend.centre_of_gravity(img->h);
    for (int a = 0; a < 3; ++a) centre[a] = c[a];
  });
}

int fw_sample_values(fw_ctx* ctx, const fw_image* img, const double* points, int64_t m, double* nearest, double* linear) {
  return guard(ctx, [&] {
    need(img, "image");
    if (m < 0) fw::invalid("sample values", "negative count");
    if (m > 0) need(points, "points");
    fw::HostVec<fw::Vec3> pts(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; ++i) pts[size_t(i)] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    fw::HostVec<double> a, b;
    ctx->backend.sample_values(img->h, pts, a, b);
    if (nearest) std::copy(a.begin(), a.end(), nearest);
    if (linear) std::copy(b.begin(), b.end(), linear);
  });
}

int fw_warp(fw_ctx* ctx, const fw_image* src, const fw_field* field, const fw_field* pre, const fw_affine* affine,
            fw_image** out) {
  return guard(ctx, [&] {
    need(src, "source");
    need(field, "field");
    need(out, "output");
    fw::AffineTransform a;
    if (affine) a = affine_from(affine);
    *out = new fw_image{ctx->backend.warp_through_field(src->h, field->h, pre ? &pre->h : nullptr, affine ? &a : nullptr)};
  });
}

int fw_warp_mask(fw_ctx* ctx, const fw_image* mask, const fw_field* field, const fw_field* pre, const fw_affine* affine,
                 fw_image* product, fw_image** out) {
  return guard(ctx, [&] {
    need(mask, "mask");
    need(field, "field");
    need(out, "output");
    fw::AffineTransform a;
    if (affine) a = affine_from(affine);
    HImage* prod = product ? &product->h : nullptr;
    fw_image* res = new fw_image{ctx->backend.warp_mask_through_field(mask->h, field->h, pre ? &pre->h : nullptr,
                                                                      affine ? &a : nullptr, prod)};
    delete product;
    *out = res;
  });
}

int fw_cc_pair(fw_ctx* ctx, const fw_image* fixed, const fw_image* moving, const fw_image* valid, uint32_t radius,
               double values[2], double norms[2], fw_field** d_fixed, fw_field** d_moving) {
  return guard(ctx, [&] {
    need(fixed, "fixed");
    need(moving, "moving");
    need(values, "values");
    need(d_fixed, "output");
    need(d_moving, "output");
    auto r = ctx->backend.cc_pair(fixed->h, moving->h, valid ? &valid->h : nullptr, radius, norms != nullptr);
    values[0] = r.values[0];
    values[1] = r.values[1];
    if (norms) {
      norms[0] = r.norms[0];
      norms[1] = r.norms[1];
    }
    *d_fixed = new fw_field{std::move(r.fields[0])};
    *d_moving = new fw_field{std::move(r.fields[1])};
  });
}

int fw_mi_dense_pair(fw_ctx* ctx, const fw_image* fixed, const fw_image* moving, const fw_image* fixed_mask,
                     const fw_image* moving_mask, uint32_t bins, double values[2], double norms[2], fw_field** d_fixed,
                     fw_field** d_moving) {
  return guard(ctx, [&] {
    need(fixed, "fixed");
    need(moving, "moving");
    need(values, "values");
    need(d_fixed, "output");
    need(d_moving, "output");
    auto r = ctx->backend.mi_pair(fixed->h, moving->h, fixed_mask ? &fixed_mask->h : nullptr,
                                  moving_mask ? &moving_mask->h : nullptr, bins, norms != nullptr);
    values[0] = r.values[0];
    values[1] = r.values[1];
    if (norms) {
      norms[0] = r.norms[0];
      norms[1] = r.norms[1];
    }
    *d_fixed = new fw_field{std::move(r.fields[0])};
    *d_moving = new fw_field{std::move(r.fields[1])};
  });
}

int fw_mi_level_create(fw_ctx* ctx, const fw_image* moving, const double* points, const uint32_t* fixed_bin, int64_t m,
                       const fw_mi_binning* binning, fw_mi_level** out) {
  return guard(ctx, [&] {
    need(moving, "moving");
    need(binning, "binning");
    need(out, "output");
    if (m < 0) fw::invalid("mattes", "negative sample count");
    if (m > 0) {
      need(points, "points");
      need(fixed_bin, "fixed bins");
    }
    fw::HostVec<fw::Vec3> pts(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; ++i) pts[size_t(i)] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    fw::HostVec<uint32_t> fb(fixed_bin, fixed_bin + m);
    fw::MiBinning b;
    b.bins = size_t(binning->bins);
    b.fixed_min = binning->fixed_min;
    b.fixed_max = binning->fixed_max;
    b.moving_min = binning->moving_min;
    b.moving_max = binning->moving_max;
    b.fixed_bin_size = binning->fixed_bin_size;
    b.moving_bin_size = binning->moving_bin_size;
    b.fixed_normalized_min = binning->fixed_normalized_min;
    b.moving_normalized_min = binning->moving_normalized_min;
    *out = new fw_mi_level{ctx, fw::hip::MiLevel(ctx->backend, moving->h, pts, fb, b)};
  });
}

static fw::LinearMovingTransform linear_from(const fw_linear_transform* t) {
  if (!t) fw::invalid("mattes", "null transform");
  fw::LinearMovingTransform lt;
  lt.outer = affine_from(&t->outer);
  lt.optimised.kind = t->kind == 0 ? fw::ParamKind::Euler3d : fw::ParamKind::Affine;
  for (int i = 0; i < 12; ++i) lt.optimised.params[i] = t->parameters[i];
  lt.optimised.center = {t->center[0], t->center[1], t->center[2]};
  return lt;
}

int fw_mi_level_value(fw_mi_level* level, const fw_linear_transform* t, double* value) {
  if (!level) return FW_ERR_INVALID;
  return guard(level->ctx, [&] {
    need(value, "output");
    *value = level->level.value(linear_from(t));
  });
}

int fw_mi_level_histogram(fw_mi_level* level, const fw_linear_transform* t, double* joint, double* fixed_marginal,
                          int64_t* valid) {
  if (!level) return FW_ERR_INVALID;
  return guard(level->ctx, [&] {
    need(joint, "joint output");
    need(fixed_marginal, "marginal output");
    need(valid, "valid output");
    std::vector<double> j, m;
    uint64_t v = 0;
    level->level.histogram_cells(linear_from(t), j, m, v);
    std::memcpy(joint, j.data(), j.size() * sizeof(double));
    std::memcpy(fixed_marginal, m.data(), m.size() * sizeof(double));
    *valid = int64_t(v);
  });
}

int fw_mi_level_values(fw_mi_level* level, const fw_linear_transform* transforms, int64_t count, double* values) {
  if (!level) return FW_ERR_INVALID;
  return guard(level->ctx, [&] {
    if (count < 0) fw::invalid("mi values", "negative count");
    if (count == 0) return;
    need(transforms, "transforms");
    need(values, "output");
    std::vector<fw::LinearMovingTransform> ts;
    for (int64_t i = 0; i < count; ++i) ts.push_back(linear_from(&transforms[i]));
    const std::vector<double> v = level->level.values(ts);
    std::copy(v.begin(), v.end(), values);
  });
}

int fw_mi_level_value_and_derivative(fw_mi_level* level, const fw_linear_transform* t, double* value,
                                     double* derivative) {
  if (!level) return FW_ERR_INVALID;
  return guard(level->ctx, [&] {
    need(value, "output");
    need(derivative, "output");
    auto r = level->level.value_and_derivative(linear_from(t));
    *value = r.first;
    std::memcpy(derivative, r.second.data(), r.second.size() * sizeof(double));
  });
}

void fw_mi_level_free(fw_mi_level* level) { delete level; }

int fw_smooth_field(fw_ctx* ctx, fw_field* field, double variance, fw_field** out) {
  return guard(ctx, [&] {
    need(field, "field");
    need(out, "output");
    HField f = std::move(field->h);
    delete field;
    *out = new fw_field{ctx->backend.smooth_field(std::move(f), variance)};
  });
}

int fw_max_scaled_norm_pair(fw_ctx* ctx, const fw_field* a, const fw_field* b, double out[2]) {
  return guard(ctx, [&] {
    need(a, "field");
    need(out, "output");
    auto r = ctx->backend.max_scaled_norm_pair(a->h, b ? &b->h : nullptr);
    out[0] = r.first;
    if (b) out[1] = r.second;
  });
}

int fw_l2_norm(fw_ctx* ctx, const fw_field* field, double* out) {
  return guard(ctx, [&] {
    need(field, "field");
    need(out, "output");
    *out = ctx->backend.l2_norm(field->h);
  });
}

int fw_scale(fw_ctx* ctx, fw_field* field, double s) {
  return guard(ctx, [&] {
    need(field, "field");
    ctx->backend.scale(field->h, s);
  });
}

int fw_axpy(fw_ctx* ctx, fw_field* acc, double a, const fw_field* x) {
  return guard(ctx, [&] {
    need(acc, "accumulator");
    need(x, "field");
    ctx->backend.axpy(acc->h, a, x->h);
  });
}

int fw_compose(fw_ctx* ctx, const fw_field* displacement, const fw_field* warping, fw_field** out) {
  return guard(ctx, [&] {
    need(displacement, "displacement");
    need(warping, "warping");
    need(out, "output");
    *out = new fw_field{ctx->backend.compose(displacement->h, warping->h)};
  });
}

static fw::hip::InvertParams params_from(const fw_invert_params* p) {
  fw::hip::InvertParams q;
  if (p) {
    q.max_iterations = p->max_iterations;
    q.mean_error_tolerance = p->mean_error_tolerance;
    q.max_error_tolerance = p->max_error_tolerance;
    q.enforce_boundary_condition = p->enforce_boundary_condition != 0;
  }
  return q;
}

static void stats_to(const fw::hip::InvertStats& s, fw_invert_stats* out) {
  if (!out) return;
  out->iterations = s.iterations;
  out->mean_error_norm = s.mean_error_norm;
  out->max_error_norm = s.max_error_norm;
}

int fw_invert(fw_ctx* ctx, const fw_field* field, fw_field* initial, const fw_invert_params* params, fw_field** out,
              fw_invert_stats* stats) {
  return guard(ctx, [&] {
    need(field, "field");
    need(initial, "initial");
    need(out, "output");
    HField init = std::move(initial->h);
    delete initial;
    auto r = ctx->backend.invert(field->h, std::move(init), params_from(params));
    stats_to(r.second, stats);
    *out = new fw_field{std::move(r.first)};
  });
}

int fw_invert_pair(fw_ctx* ctx, fw_field* field, fw_field* inverse, const fw_invert_params* params,
                   fw_field** new_inverse, fw_invert_stats* inverse_stats, fw_field** new_field,
                   fw_invert_stats* field_stats) {
  return guard(ctx, [&] {
    need(field, "field");
    need(inverse, "inverse");
    need(new_inverse, "output");
    need(new_field, "output");
    HField f = std::move(field->h);
    HField inv = std::move(inverse->h);
    delete field;
    delete inverse;
    const auto p = params_from(params);
    auto a = ctx->backend.invert(f, std::move(inv), p);
    auto b = ctx->backend.invert(a.first, std::move(f), p);
    stats_to(a.second, inverse_stats);
    stats_to(b.second, field_stats);
    *new_inverse = new fw_field{std::move(a.first)};
    *new_field = new fw_field{std::move(b.first)};
  });
}

int fw_resample_field(fw_ctx* ctx, const fw_field* field, const fw_grid* grid, fw_field** out) {
  return guard(ctx, [&] {
    need(field, "field");
    need(out, "output");
    *out = new fw_field{ctx->backend.resample_field(field->h, grid_from(grid))};
  });
}

}  // extern "C"

// ------------------------------------------------------------------ whole registrations ---

extern "C" {

int fw_set_profiling(fw_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->backend.ctx().set_profiling(on != 0); });
}

int fw_set_profiling_focus(fw_ctx* ctx, const char* family, int every) {
  return guard(ctx, [&] { ctx->backend.ctx().set_profiling_focus(family ? std::string(family) : std::string(), every); });
}

int fw_take_profile(fw_ctx* ctx, char* text, size_t capacity) {
  return guard(ctx, [&] {
    need(text, "output");
    const std::string p = ctx->backend.ctx().take_profile();
    if (capacity == 0) fw::invalid("profile", "zero capacity");
    const size_t n = std::min(p.size(), capacity - 1);
    std::memcpy(text, p.data(), n);
    text[n] = '\0';
  });
}

static fw::CapacityPlan capacity_of(const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid) {
  need(plan, "plan");
  need(fixed_grid, "fixed grid");
  need(moving_grid, "moving grid");
  const fw::RegisterPlan p = fw::plan_from(plan);
  return fw::plan_capacity(p, fw::grid_of(fixed_grid), fw::grid_of(moving_grid), size_t(std::max(plan->channels, 0)));
}

static void fill(const fw::CapacityPlan& c, fw_capacity* out) {
  out->fixed_voxels = c.fixed_voxels;
  out->moving_voxels = c.moving_voxels;
  out->smoothing_cache_bytes = c.smoothing_cache_bytes;
  out->linear_bytes = c.linear_bytes;
  out->syn_bytes = c.syn_bytes;
  out->peak_bytes = c.peak_bytes;
  out->largest_buffer_bytes = c.largest_buffer_bytes;
}

int fw_plan_capacity(const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid, fw_capacity* out) {
  try {
    need(out, "output");
    fill(capacity_of(plan, fixed_grid, moving_grid), out);
    return FW_OK;
  } catch (const std::exception&) {
    return FW_ERR_INVALID;
  }
}

int fw_memory_budget(fw_ctx* ctx, uint64_t* bytes) {
  return guard(ctx, [&] {
    need(bytes, "output");
    *bytes = ctx->backend.memory_budget();
  });
}

int fw_fits(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid, int32_t* fits) {
  return guard(ctx, [&] {
    need(fits, "output");
    *fits = capacity_of(plan, fixed_grid, moving_grid).peak_bytes <= ctx->backend.memory_budget() ? 1 : 0;
  });
}

int fw_registrations_at_once(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid,
                             uint32_t max, uint32_t* count) {
  return guard(ctx, [&] {
    need(count, "output");
    *count = fw::registrations_at_once(capacity_of(plan, fixed_grid, moving_grid), ctx->backend.memory_budget(), max);
  });
}

int fw_write_affine_mat(const char* path, const fw_affine* affine) {
  try {
    need(path, "path");
    need(affine, "affine");
    fw::write_affine_mat(path, affine_from(affine));
    return FW_OK;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return FW_ERR_INVALID;
  }
}

int fw_read_affine_mat(const char* path, fw_affine* out) {
  try {
    need(path, "path");
    need(out, "output");
    const fw::AffineTransform t = fw::read_affine_mat(path);
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) out->matrix[3 * r + c] = t.matrix[r][c];
      out->translation[r] = t.translation[r];
      out->center[r] = t.center[r];
    }
    return FW_OK;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return FW_ERR_INVALID;
  }
}

}  // extern "C"

namespace {

void affine_to(const fw::AffineTransform& t, fw_affine* out) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) out->matrix[3 * r + c] = t.matrix[r][c];
    out->translation[r] = t.translation[r];
    out->center[r] = t.center[r];
  }
}

std::vector<float> field_values(const Grid& g, const double* v) {
  need(v, "field values");
  std::vector<float> out(3 * g.num_voxels());
  for (size_t i = 0; i < out.size(); ++i) out[i] = float(v[i]);
  return out;
}

template <typename F>
int host_guard(F&& f) {
  try {
    f();
    return FW_OK;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return FW_ERR_INVALID;
  }
}

}  // namespace

extern "C" {

int fw_write_composite_h5(const char* path, const fw_chain_element* elements, uint32_t count) {
  return host_guard([&] {
    need(path, "path");
    if (count == 0) fw::invalid("ITK HDF5 transform", "cannot write an empty chain");
    need(elements, "elements");
    std::vector<fw::ChainElement> chain(count);
    for (uint32_t i = 0; i < count; ++i) {
      const fw_chain_element& e = elements[i];
      if (e.kind == FW_CHAIN_AFFINE) {
        chain[i].affine = affine_from(&e.affine);
      } else if (e.kind == FW_CHAIN_FIELD) {
        chain[i].is_field = true;
        chain[i].grid = grid_from(&e.grid);
        chain[i].field = field_values(chain[i].grid, e.values);
      } else {
        fw::invalid("ITK HDF5 transform", "unknown chain element kind " + std::to_string(e.kind));
      }
    }
    fw::write_composite_h5(path, chain);
  });
}

int fw_read_composite_h5(const char* path, fw_chain_element* elements, uint32_t capacity, uint32_t* count) {
  return host_guard([&] {
    need(path, "path");
    need(count, "count");
    if (capacity > 0) need(elements, "elements");
    const std::vector<fw::ChainElement> chain = fw::read_composite_h5(path);
    *count = uint32_t(chain.size());
    for (size_t i = 0; i < chain.size() && i < capacity; ++i) {
      fw_chain_element& out = elements[i];
      const fw::ChainElement& e = chain[i];
      out.kind = e.is_field ? FW_CHAIN_FIELD : FW_CHAIN_AFFINE;
      if (!e.is_field) {
        affine_to(e.affine, &out.affine);
        continue;
      }
      grid_to(e.grid, &out.grid);
      if (out.values)
        for (size_t k = 0; k < e.field.size(); ++k) out.values[k] = double(e.field[k]);
    }
  });
}

int fw_write_composite_pair(const char* prefix, const fw_affine* affine, const fw_grid* grid, const double* field,
                            const double* inverse_field) {
  return host_guard([&] {
    need(prefix, "prefix");
    need(affine, "affine");
    if (!grid) {
      fw::write_composite_pair(prefix, affine_from(affine), nullptr, {}, {});
      return;
    }
    const Grid g = grid_from(grid);
    fw::write_composite_pair(prefix, affine_from(affine), &g, field_values(g, field), field_values(g, inverse_field));
  });
}

int fw_write_composite_fields(fw_ctx* ctx, const char* prefix, const fw_affine* affine, const fw_field* forward,
                              const fw_field* inverse) {
  return guard(ctx, [&] {
    need(prefix, "prefix");
    need(affine, "affine");
    need(forward, "forward field");
    need(inverse, "inverse field");
    if (!(forward->h.grid == inverse->h.grid)) fw::invalid("ITK HDF5 transform", "the two fields are on different grids");
    const auto as_f32 = [&](const fw_field* f) {
      const fw::VectorField v = ctx->backend.download_field(f->h);
      std::vector<float> out(3 * v.data.size());
      for (size_t i = 0; i < v.data.size(); ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + size_t(c)] = float(v.data[i][c]);
      return out;
    };
    fw::write_composite_pair(prefix, affine_from(affine), &forward->h.grid, as_f32(forward), as_f32(inverse));
  });
}

void fw_plan_cohort(fw_plan* plan, int32_t channels) {
  if (plan) fw::cohort_plan(plan, channels);
}

int32_t fw_last_trace(fw_trace_step* out, int32_t capacity) { return fw::copy_last_trace(out, capacity); }

int fw_register_resident(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, const fw_image* const* moving,
                         fw_field** forward, fw_field** inverse, fw_result* result, const volatile int32_t* cancel) {
  return guard(ctx, [&] {
    need(plan, "plan");
    if (!fixed || !moving) fw::invalid("register", "null image list");
    const fw::RegisterPlan p = fw::plan_from(plan);
    std::vector<const HImage*> f, m;
    for (int32_t c = 0; c < plan->channels; ++c) {
      if (!fixed[c] || !moving[c]) fw::invalid("register", "null image");
      f.push_back(&fixed[c]->h);
      m.push_back(&moving[c]->h);
    }
    fw::hip::HipLinearEngine engine(ctx->backend);
    std::optional<HField> fwd, inv;
    const fw::RegisterOutput o = fw::register_resident(ctx->backend, engine, p, f, m, cancel, forward ? &fwd : nullptr,
                                                       inverse ? &inv : nullptr);
    fw::result_to(o, result);
    if (forward) *forward = fwd ? new fw_field{std::move(*fwd)} : nullptr;
    if (inverse) *inverse = inv ? new fw_field{std::move(*inv)} : nullptr;
  });
}

int fw_syn_stage(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, const fw_image* const* moving,
                 const fw_affine* initial, const fw_image* fixed_mask, const fw_image* moving_mask, const fw_field* prior,
                 const fw_field* prior_inverse, fw_field** forward, fw_field** inverse, fw_result* result,
                 const volatile int32_t* cancel) {
  return guard(ctx, [&] {
    need(plan, "plan");
    if (!fixed || !moving) fw::invalid("syn stage", "null image list");
    std::vector<const HImage*> f, m;
    for (int32_t c = 0; c < plan->channels; ++c) {
      if (!fixed[c] || !moving[c]) fw::invalid("syn stage", "null image");
      f.push_back(&fixed[c]->h);
      m.push_back(&moving[c]->h);
    }
    fw::SynStageInputs<HipBackend> extra;
    extra.fixed_mask = fixed_mask ? &fixed_mask->h : nullptr;
    extra.moving_mask = moving_mask ? &moving_mask->h : nullptr;
    extra.prior = prior ? &prior->h : nullptr;
    extra.prior_inverse = prior_inverse ? &prior_inverse->h : nullptr;
    std::optional<HField> fwd, inv;
    const fw::RegisterOutput o = fw::run_syn_stage(ctx->backend, plan, f, m, initial, extra, cancel, fw::FieldOutputs{},
                                                   forward ? &fwd : nullptr, inverse ? &inv : nullptr);
    fw::result_to(o, result);
    if (forward) *forward = fwd ? new fw_field{std::move(*fwd)} : nullptr;
    if (inverse) *inverse = inv ? new fw_field{std::move(*inv)} : nullptr;
  });
}

namespace {
// The per-subject records of the calling thread's last fw_register_batch.
std::vector<std::vector<fw_trace_step>>& last_batch_traces() {
  static thread_local std::vector<std::vector<fw_trace_step>> t;
  return t;
}
}  // namespace

namespace {
constexpr uint32_t kMaxWorkers = 16;
std::vector<fw::hip::HipBackend*> take_workers(fw_ctx* ctx, uint32_t workers, size_t jobs) {
  if (workers > kMaxWorkers) fw::invalid("workers", "at most " + std::to_string(kMaxWorkers) + " workers");
  const size_t w = std::max<size_t>(1, std::min<size_t>(workers ? workers : 8, jobs));
  return ctx->workers.take(w);
}
}  // namespace

int fw_linear_batch(fw_ctx* ctx, const fw_plan* plan, uint32_t fits, const fw_image* const* fixed,
                    const fw_image* const* moving, uint32_t workers, fw_result* results, const volatile int32_t* cancel) {
  return guard(ctx, [&] {
    need(plan, "plan");
    if (!fixed || !moving) fw::invalid("register", "null image list");
    if (fits == 0) fw::invalid("linear batch", "no fits");
    fw::RegisterPlan p = fw::plan_from(plan);
    p.run_syn = false;
    std::vector<fw::hip::LinearFitJob> jobs(fits);
    for (uint32_t i = 0; i < fits; ++i) {
      if (!fixed[i] || !moving[i]) fw::invalid("register", "null image");
      jobs[i].fixed_raw = &fixed[i]->h;
      jobs[i].moving_raw = &moving[i]->h;
    }
    ctx->backend.ctx().sync();  // the images are complete before other streams read them
    fw::hip::run_linear_fits(take_workers(ctx, workers, fits), p, jobs, cancel);
    auto& traces = last_batch_traces();
    traces.assign(fits, {});
    for (uint32_t i = 0; i < fits; ++i) {
      fw::RegisterOutput o;
      o.linear = jobs[i].out;
      o.initial = jobs[i].out.initial;
      o.total = jobs[i].out.total;
      o.seconds_preprocess = jobs[i].seconds_preprocess;
      o.seconds_linear = jobs[i].seconds_linear;
      o.seconds_total = o.seconds_preprocess + o.seconds_linear;
      fw::result_to(o, results ? &results[i] : nullptr);
      traces[i] = fw::last_trace();
    }
  });
}

int fw_register_batch(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, uint32_t subjects,
                      const fw_image* const* moving, uint32_t workers, fw_field** forward, fw_field** inverse,
                      fw_result* results, const volatile int32_t* cancel) {
  return guard(ctx, [&] {
    need(plan, "plan");
    if (!fixed || !moving) fw::invalid("register", "null image list");
    if (subjects == 0) fw::invalid("register batch", "no subjects");
    const fw::RegisterPlan p = fw::plan_from(plan);
    const size_t nc = size_t(plan->channels);
    std::vector<const HImage*> f;
    std::vector<std::vector<const HImage*>> m(subjects);
    for (size_t c = 0; c < nc; ++c) {
      if (!fixed[c]) fw::invalid("register", "null image");
      f.push_back(&fixed[c]->h);
    }
    for (size_t s = 0; s < subjects; ++s)
      for (size_t c = 0; c < nc; ++c) {
        const fw_image* im = moving[s * nc + c];
        if (!im) fw::invalid("register", "null image");
        m[s].push_back(&im->h);
      }
    std::vector<std::optional<HField>> fwd, inv;
    const std::vector<fw::RegisterOutput> o =
        fw::hip::register_batch(ctx->backend, take_workers(ctx, workers, subjects), p, f, m, cancel,
                                forward ? &fwd : nullptr, inverse ? &inv : nullptr);
    auto& traces = last_batch_traces();
    traces.assign(subjects, {});
    for (size_t s = 0; s < subjects; ++s) {
      fw::result_to(o[s], results ? &results[s] : nullptr);
      traces[s] = fw::last_trace();
      if (forward) forward[s] = fwd[s] ? new fw_field{std::move(*fwd[s])} : nullptr;
      if (inverse) inverse[s] = inv[s] ? new fw_field{std::move(*inv[s])} : nullptr;
    }
  });
}

int32_t fw_last_batch_trace(uint32_t subject, fw_trace_step* out, int32_t capacity) {
  const auto& all = last_batch_traces();
  if (subject >= all.size()) return -1;
  const auto& t = all[subject];
  const size_t n = out && capacity > 0 ? std::min(t.size(), size_t(capacity)) : 0;
  for (size_t i = 0; i < n; ++i) out[i] = t[i];
  return int32_t(t.size());
}

int fw_register(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const double* const* fixed,
                const fw_grid* moving_grid, const double* const* moving, double* forward, double* inverse,
                fw_result* result, const volatile int32_t* cancel) {
  return guard(ctx, [&] {
    fw::hip::HipLinearEngine engine(ctx->backend);
    fw::run_register(ctx->backend, engine, plan, fixed_grid, fixed, moving_grid, moving, forward, inverse, result, cancel);
  });
}

}  // extern "C"
