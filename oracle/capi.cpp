// ORACLE — test infrastructure only (see oracle.hpp). A flat C interface over the f64 restatement
// so the Python tests can run the same inputs through the oracle and through the HIP library.
// Grids use the fw_grid struct of include/fastwarp_hip.h; images are f64[n], fields f64[3n]
// interleaved (dx, dy, dz).
#include <cstring>
#include <string>

#include "../include/fastwarp_hip.h"
#include "../fastwarp_amd/csrc/plan.hpp"
#include "engine.hpp"
#include "linear.hpp"
#include "syn.hpp"
#include "oracle.hpp"

using namespace fw;
using namespace fw::oracle;

namespace {

thread_local std::string g_error;

Grid grid_from(const fw_grid* g) {
  Mat3 dir;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dir[r][c] = g->direction[3 * r + c];
  return Grid({size_t(g->size[0]), size_t(g->size[1]), size_t(g->size[2])}, {g->spacing[0], g->spacing[1], g->spacing[2]},
              {g->origin[0], g->origin[1], g->origin[2]}, dir);
}

Image image_from(const fw_grid* g, const double* data) {
  Grid gr = grid_from(g);
  return Image(gr, std::vector<double>(data, data + gr.num_voxels()));
}

VectorField field_from(const fw_grid* g, const double* aos) {
  Grid gr = grid_from(g);
  std::vector<Vec3> d(gr.num_voxels());
  for (size_t i = 0; i < d.size(); ++i) d[i] = {aos[3 * i], aos[3 * i + 1], aos[3 * i + 2]};
  return VectorField(gr, std::move(d));
}

void field_to(const VectorField& f, double* aos) {
  for (size_t i = 0; i < f.data.size(); ++i)
    for (int c = 0; c < 3; ++c) aos[3 * i + c] = f.data[i][c];
}

AffineTransform affine_from(const fw_affine* a) {
  AffineTransform t;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.matrix[r][c] = a->matrix[3 * r + c];
    t.translation[r] = a->translation[r];
    t.center[r] = a->center[r];
  }
  return t;
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // namespace

extern "C" {

const char* fwo_last_error() { return g_error.c_str(); }

// pyramid.rs:429-437 `smooth_recursive_gaussian` in f64 (the drivers' host restatement).
int fwo_smooth_image(const fw_grid* grid, const double* data, const double sigma_mm[3], double* out) {
  return guard([&] {
    Image v = image_from(grid, data);
    Image r = fw::oracle::smooth_recursive_gaussian(v, {sigma_mm[0], sigma_mm[1], sigma_mm[2]});
    std::copy(r.data.begin(), r.data.end(), out);
  });
}

// preprocess.rs:159-232 on the host in f64 (the host code both engines share).
int fwo_preprocess(const fw_grid* grid, const double* data, float lower, float upper, double* out) {
  return guard([&] {
    Image v = image_from(grid, data);
    Image r = fw::preprocess_volume(v, lower, upper);
    std::copy(r.data.begin(), r.data.end(), out);
  });
}

// register.rs:112-142 on the host in f64: the isotropic grid into *out_grid and, when `out` holds
// at least its voxel count (`capacity`), the resampled values.
int fwo_resample_isotropic(const fw_grid* grid, const double* data, double mm, fw_grid* out_grid, double* out,
                           int64_t capacity) {
  return guard([&] {
    Image v = image_from(grid, data);
    Image r = fw::oracle::resample_isotropic(v, mm);
    for (int a = 0; a < 3; ++a) {
      out_grid->size[a] = int64_t(r.grid.size[a]);
      out_grid->spacing[a] = r.grid.spacing[a];
      out_grid->origin[a] = r.grid.origin[a];
      for (int c = 0; c < 3; ++c) out_grid->direction[3 * a + c] = r.grid.direction[a][c];
    }
    if (out) {
      if (capacity < int64_t(r.data.size())) fw::invalid("register", "output buffer too small");
      std::copy(r.data.begin(), r.data.end(), out);
    }
  });
}

// init.rs:47-86 on the host in f64.
int fwo_centre_of_gravity(const fw_grid* grid, const double* data, double* centre) {
  return guard([&] {
    Image v = image_from(grid, data);
    const fw::Vec3 c = fw::centre_of_gravity(v);
    for (int a = 0; a < 3; ++a) centre[a] = c[a];
  });
}

// mi.rs:161-169 nearest-voxel value and interp.rs:43-87 linear value at physical points, NaN outside.
int fwo_sample_values(const fw_grid* grid, const double* data, const double* points, int64_t m, double* nearest,
                      double* linear) {
  return guard([&] {
    Image v = image_from(grid, data);
    for (int64_t i = 0; i < m; ++i) {
      const fw::Vec3 p{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
      const auto q = oracle::nearest(v.grid, p);
      nearest[i] = q ? v.data[*q] : std::nan("");
      const auto l = oracle::lerp_point(v, p);
      linear[i] = l ? *l : std::nan("");
    }
  });
}

int fwo_warp(const fw_grid* src_grid, const double* src, const fw_grid* field_grid, const double* field,
             const fw_grid* pre_grid, const double* pre, const fw_affine* affine, int nearest, const double* product,
             double* out) {
  return guard([&] {
    Image s = image_from(src_grid, src);
    VectorField f = field_from(field_grid, field);
    VectorField p;
    if (pre) p = field_from(pre_grid, pre);
    AffineTransform a;
    if (affine) a = affine_from(affine);
    Image r;
    if (nearest) {
      Image prod;
      if (product) prod = image_from(field_grid, product);
      r = warp_mask_through_field(s, f, pre ? &p : nullptr, affine ? &a : nullptr, product ? &prod : nullptr);
    } else {
      r = warp_through_field(s, f, pre ? &p : nullptr, affine ? &a : nullptr);
    }
    std::memcpy(out, r.data.data(), r.data.size() * sizeof(double));
  });
}

// CC on a virtual grid; moving_affine (may be null) maps virtual points to moving points.
int fwo_cc(const fw_grid* grid, const double* fixed, const fw_grid* moving_grid, const double* moving,
           const fw_affine* moving_affine, const double* mask, int radius, int want_derivative, double* value,
           int64_t* num_valid, double* local_cc, double* coefficient, double* derivative, uint8_t* valid) {
  return guard([&] {
    Grid g = grid_from(grid);
    Image f = image_from(grid, fixed);
    Image m = image_from(moving_grid, moving);
    Image mk;
    if (mask) mk = image_from(grid, mask);
    AffineTransform a;
    PointMap id, mm;
    if (moving_affine) {
      a = affine_from(moving_affine);
      mm.affine = &a;
    }
    CcEvaluation e = cc_evaluate(g, f, m, id, mm, mask ? &mk : nullptr, size_t(radius), want_derivative != 0);
    *value = e.value;
    if (num_valid) *num_valid = int64_t(e.num_valid);
    const size_t n = g.num_voxels();
    if (local_cc) std::memcpy(local_cc, e.local_cc.data(), n * sizeof(double));
    if (coefficient) std::memcpy(coefficient, e.coefficient.data(), n * sizeof(double));
    if (valid) std::memcpy(valid, e.valid.data(), n);
    if (derivative && want_derivative)
      for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) derivative[3 * i + c] = e.derivative[i][c];
  });
}

// Dense Mattes MI on the fixed grid as virtual grid; `field` (may be null) is an f32
// displacement-field moving transform on `field_grid` (3 floats per voxel interleaved).
int fwo_mi_dense(const fw_grid* fixed_grid, const double* fixed, const fw_grid* moving_grid, const double* moving,
                 const fw_grid* field_grid, const float* field, const double* fixed_mask, const double* moving_mask,
                 int bins, int want_derivative, double* value, int64_t* num_valid, double* derivative) {
  return guard([&] {
    Grid g = grid_from(fixed_grid);
    Image f = image_from(fixed_grid, fixed);
    Image m = image_from(moving_grid, moving);
    Image fm, mm;
    if (fixed_mask) fm = image_from(fixed_grid, fixed_mask);
    if (moving_mask) mm = image_from(moving_grid, moving_mask);
    DisplacementTransform dt;
    PointMap id, mmap;
    if (field) {
      dt.grid = grid_from(field_grid);
      dt.data.assign(field, field + 3 * dt.grid.num_voxels());
      mmap.field = &dt;
    }
    MiDenseEvaluation e = mi_dense(g, f, m, id, mmap, fixed_mask ? &fm : nullptr, moving_mask ? &mm : nullptr, size_t(bins),
                                   want_derivative != 0);
    *value = e.value;
    if (num_valid) *num_valid = int64_t(e.num_valid);
    if (derivative && want_derivative)
      for (size_t i = 0; i < e.derivative.size(); ++i)
        for (int c = 0; c < 3; ++c) derivative[3 * i + c] = e.derivative[i][c];
  });
}

// Sampled (linear-stage) Mattes MI with prepared fixed bins.
// mi.rs:702-769 on given cells: the value and (if ratio is not null) the ratio weights.
int fwo_mi_finish(int64_t bins, double moving_bin_size, int64_t num_valid, const double* joint, const double* fixed_marginal,
                  double* value, double* ratio) {
  return guard([&] {
    const size_t nb = size_t(bins);
    HistogramResult r = fw::oracle::finish_histogram(nb, moving_bin_size, size_t(num_valid), std::vector<double>(joint, joint + nb * nb),
                                         std::vector<double>(fixed_marginal, fixed_marginal + nb), ratio != nullptr);
    *value = r.value;
    if (ratio)
      for (size_t i = 0; i < nb * nb; ++i) ratio[i] = r.ratio[i];
  });
}

int fwo_mi_sampled(const fw_mi_binning* b, const fw_grid* moving_grid, const double* moving, const double* points,
                   const uint32_t* fixed_bin, int64_t m, const fw_linear_transform* t, int want_derivative, double* value,
                   double* derivative, int64_t* num_valid, double* joint_pdf) {
  return guard([&] {
    Binning bn;
    bn.bins = size_t(b->bins);
    bn.fixed_min = b->fixed_min;
    bn.fixed_max = b->fixed_max;
    bn.moving_min = b->moving_min;
    bn.moving_max = b->moving_max;
    bn.fixed_bin_size = b->fixed_bin_size;
    bn.moving_bin_size = b->moving_bin_size;
    bn.fixed_normalized_min = b->fixed_normalized_min;
    bn.moving_normalized_min = b->moving_normalized_min;
    Image mv = image_from(moving_grid, moving);
    std::vector<Vec3> pts(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; ++i) pts[size_t(i)] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    std::vector<uint32_t> fb(fixed_bin, fixed_bin + m);
    LinearMovingTransform lt;
    lt.outer = affine_from(&t->outer);
    lt.optimised.kind = t->kind == 0 ? ParamKind::Euler3d : ParamKind::Affine;
    for (int i = 0; i < 12; ++i) lt.optimised.params[i] = t->parameters[i];
    lt.optimised.center = {t->center[0], t->center[1], t->center[2]};
    MiSampledEvaluation e = mi_sampled(bn, mv, pts, fb, lt, want_derivative != 0);
    *value = e.value;
    if (num_valid) *num_valid = int64_t(e.num_valid);
    if (derivative && want_derivative) std::memcpy(derivative, e.derivative.data(), e.derivative.size() * sizeof(double));
    if (joint_pdf) std::memcpy(joint_pdf, e.hist.joint_pdf.data(), e.hist.joint_pdf.size() * sizeof(double));
  });
}

int fwo_gaussian_taps(double variance, double max_error, int64_t max_width, double* out, int64_t capacity, int64_t* len) {
  return guard([&] {
    auto t = gaussian_taps(variance, max_error, size_t(max_width));
    *len = int64_t(t.size());
    if (int64_t(t.size()) <= capacity) std::memcpy(out, t.data(), t.size() * sizeof(double));
  });
}

double fwo_bessel(int n, double y) { return n == 0 ? bessel_i0(y) : n == 1 ? bessel_i1(y) : bessel_i(n, y); }

int fwo_smooth(const fw_grid* grid, const double* field, double variance, double* out) {
  return guard([&] { field_to(smooth_field(field_from(grid, field), variance), out); });
}

int fwo_compose(const fw_grid* grid, const double* displacement, const double* warping, double* out) {
  return guard([&] { field_to(compose(field_from(grid, displacement), field_from(grid, warping)), out); });
}

int fwo_invert(const fw_grid* grid, const double* field, const double* initial, const fw_invert_params* p, double* out,
               fw_invert_stats* stats) {
  return guard([&] {
    VectorField f = field_from(grid, field);
    VectorField init;
    if (initial) init = field_from(grid, initial);
    oracle::InvertParams q;
    q.max_iterations = p->max_iterations;
    q.mean_error_tolerance = p->mean_error_tolerance;
    q.max_error_tolerance = p->max_error_tolerance;
    q.enforce_boundary_condition = p->enforce_boundary_condition != 0;
    auto r = invert(f, initial ? &init : nullptr, q);
    field_to(r.first, out);
    stats->iterations = r.second.iterations;
    stats->mean_error_norm = r.second.mean_error_norm;
    stats->max_error_norm = r.second.max_error_norm;
  });
}

int fwo_resample(const fw_grid* grid, const double* field, const fw_grid* target, double* out) {
  return guard([&] { field_to(resample_field(field_from(grid, field), grid_from(target)), out); });
}

int fwo_norms(const fw_grid* grid, const double* field, double* max_scaled, double* l2) {
  return guard([&] {
    VectorField f = field_from(grid, field);
    *max_scaled = max_scaled_norm(f);
    *l2 = l2_norm(f);
  });
}

double fwo_cubic_bspline(double u, int derivative) { return derivative ? cubic_bspline_derivative(u) : cubic_bspline(u); }

}  // extern "C"

// The whole CPU engine: the generic drivers over the oracle's f64 ops (engine="cpu").
extern "C" int32_t fwo_last_trace(fw_trace_step* out, int32_t capacity) { return fw::copy_last_trace(out, capacity); }

namespace {
InversionScript g_script;
bool g_script_on = false;
}  // namespace

extern "C" int fwo_register(const fw_plan* plan, const fw_grid* fixed_grid, const double* const* fixed,
                            const fw_grid* moving_grid, const double* const* moving, double* forward, double* inverse,
                            fw_result* result) {
  return guard([&] {
    CpuBackend backend;
    CpuLinearEngine engine;
    if (g_script_on) {
      g_script.next = 0;
      g_script.disagreements.clear();
      g_script.disagreement_calls.clear();
      backend.script = &g_script;
    }
    fw::run_register(backend, engine, plan, fixed_grid, fixed, moving_grid, moving, forward, inverse, result, nullptr);
    if (g_script_on && g_script.next != g_script.steps.size())
      invalid("lockstep", "the run made fewer inversions than the script holds");
  });
}

// fw_syn_stage on the CPU engine: f64 host images (each channel's fixed on fixed_grid, moving on
// moving_grid), masks on their own grids (data may be null), earlier fields (3n interleaved on
// prior_grid, may be null), forward / inverse 3n on the fixed grid (may be null).
extern "C" int fwo_syn_stage(const fw_plan* plan, const fw_grid* fixed_grid, const double* const* fixed,
                             const fw_grid* moving_grid, const double* const* moving, const fw_affine* initial,
                             const fw_grid* fixed_mask_grid, const double* fixed_mask, const fw_grid* moving_mask_grid,
                             const double* moving_mask, const fw_grid* prior_grid, const double* prior,
                             const double* prior_inverse, double* forward, double* inverse, fw_result* result) {
  return guard([&] {
    if (!plan) invalid("syn stage", "null plan");
    const Grid fg = grid_of(fixed_grid), mg = grid_of(moving_grid);
    std::vector<Image> f, m;
    for (int32_t c = 0; c < plan->channels; ++c) {
      f.emplace_back(fg, std::vector<double>(fixed[c], fixed[c] + fg.num_voxels()));
      m.emplace_back(mg, std::vector<double>(moving[c], moving[c] + mg.num_voxels()));
    }
    std::vector<const Image*> fp, mp;
    for (size_t c = 0; c < f.size(); ++c) {
      fp.push_back(&f[c]);
      mp.push_back(&m[c]);
    }
    auto image = [](const fw_grid* g, const double* d) -> std::optional<Image> {
      if (!d) return std::nullopt;
      const Grid gg = grid_of(g);
      return Image(gg, std::vector<double>(d, d + gg.num_voxels()));
    };
    auto field = [](const fw_grid* g, const double* d) -> std::optional<VectorField> {
      if (!d) return std::nullopt;
      const Grid gg = grid_of(g);
      std::vector<Vec3> v(gg.num_voxels());
      for (size_t i = 0; i < v.size(); ++i) v[i] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
      return VectorField(gg, std::move(v));
    };
    const std::optional<Image> fm = image(fixed_mask_grid, fixed_mask), mm = image(moving_mask_grid, moving_mask);
    const std::optional<VectorField> pw = field(prior_grid, prior), pi = field(prior_grid, prior_inverse);
    SynStageInputs<CpuBackend> extra;
    extra.fixed_mask = fm ? &*fm : nullptr;
    extra.moving_mask = mm ? &*mm : nullptr;
    extra.prior = pw ? &*pw : nullptr;
    extra.prior_inverse = pi ? &*pi : nullptr;
    CpuBackend backend;
    const RegisterOutput o = run_syn_stage(backend, plan, fp, mp, initial, extra, nullptr, FieldOutputs{forward, inverse});
    result_to(o, result);
  });
}

// 1: the CPU engine's SyN stages run the product's generic loop over the oracle's ops; 0: its own.
extern "C" void fwo_generic_syn(int on) { use_generic_syn() = on != 0; }
// The same for the linear chain.
extern "C" void fwo_generic_linear(int on) { use_generic_linear() = on != 0; }

// The next fwo_register calls replay these inversion step counts in call order (n < 0: off).
extern "C" void fwo_lockstep(const int32_t* steps, int64_t n) {
  g_script_on = n >= 0;
  g_script.steps.assign(steps, steps + (n > 0 ? n : 0));
}

// The stop decisions of the last lockstep run where the oracle's own rule differed: per entry the
// inversion call, the body count, the rule's decision (1 continue) and the errors it decided on.
// Returns the count; copies up to `capacity`.
extern "C" int64_t fwo_lockstep_disagreements(int64_t* call, int32_t* body, int32_t* natural, double* max_error,
                                              double* mean_error, int64_t capacity) {
  const auto& d = g_script.disagreements;
  for (size_t i = 0; i < d.size() && int64_t(i) < capacity; ++i) {
    call[i] = int64_t(g_script.disagreement_calls[i]);
    body[i] = int32_t(d[i].body);
    natural[i] = d[i].natural ? 1 : 0;
    max_error[i] = d[i].max_error;
    mean_error[i] = d[i].mean_error;
  }
  return int64_t(d.size());
}
