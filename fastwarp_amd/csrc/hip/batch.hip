// The batched SyN loop: the iteration of syn.hpp (crates/fastwarp-reg/src/syn.rs:776-986) for
// several subjects at once, each step one launch over every active subject.
#include "batch.hpp"

#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <thread>

#include "geometry.hpp"

namespace fw::hip {
namespace {

using clock_type = std::chrono::steady_clock;

struct Subject {
  std::optional<HField> f2m, f2m_inv, m2m, m2m_inv;
  // the buffers the next compositions write (they swap with f2m / m2m every iteration)
  std::optional<HField> alt_f, alt_m;
  std::optional<WindowConvergenceMonitor> monitor;
  SynLevelTrace trace;
  uint32_t current = 0;
  bool active = false;
};

// smooth_field_pair over any number of fields on one grid: the same taps, pass shape and kernels,
// so each result is the bits of smooth_field_pair / smooth_field on it.
std::vector<HField> smooth_fields(HipBackend& b, std::vector<HField> in, double variance) {
  if (variance <= 0.0 || in.empty()) return in;
  DeviceContext& ctx = b.ctx();
  const Grid g = in[0].grid;
  const Dims d = dims_of(g);
  const size_t n = g.num_voxels();
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
  GaussTaps3 t{};
  int* radii[3] = {&t.rx, &t.ry, &t.rz};
  float* dst[3] = {t.x, t.y, t.z};
  bool fits = true;
  for (int q = 0; q < 3 && fits; ++q) {
    const std::vector<double>& taps = ctx.host_taps(variance, g.size[q]);
    *radii[q] = int(taps.size() / 2);
    if (taps.size() > 17) fits = false;
    else
      for (size_t k = 0; k < taps.size(); ++k) dst[q][k] = float(taps[k]);
  }
  const size_t m = in.size();
  if (!fits || !gauss3_fits(d, t) || m > size_t(kBatchFields) || 3ull * m * size_t(d.nz) > 65535) {
    // the pair form, two fields at a time (it falls back to single fields itself)
    std::vector<HField> out;
    for (size_t q = 0; q < m; q += 2) {
      if (q + 1 < m) {
        auto r = b.smooth_field_pair(std::move(in[q]), std::move(in[q + 1]), variance);
        out.push_back(std::move(r.first));
        out.push_back(std::move(r.second));
      } else {
        out.push_back(b.smooth_field(std::move(in[q]), variance));
      }
    }
    return out;
  }
  std::vector<HField> out;
  FieldPtrs p{};
  p.count = int(m);
  for (size_t q = 0; q < m; ++q) {
    out.push_back(HField{g, DevBuf(&ctx, 3 * n)});
    p.in[q] = in[q].buf.get();
    p.orig[q] = in[q].buf.get();
    p.out[q] = out[q].buf.get();
  }
  hipStream_t s = ctx.stream();
  const double fields = double(m);
  if (t.rx == t.ry && t.ry == t.rz && ctx.tuning().get("gauss.one_pass", uint64_t(d.n()), gauss3_pays(d) ? 1 : 0)) {
    Timed tm(ctx, "k_gauss3", fields * (24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0)));
    if (launch_gauss3_set(p, d, t, float(w1), float(w2), s)) return out;
  }
  std::vector<DevBuf> tmp;
  FieldPtrs xy{}, z{};
  xy.count = z.count = int(m);
  for (size_t q = 0; q < m; ++q) {
    tmp.emplace_back(&ctx, 3 * n);
    xy.in[q] = in[q].buf.get();
    xy.out[q] = tmp[q].get();
    z.in[q] = tmp[q].get();
    z.orig[q] = in[q].buf.get();
    z.out[q] = out[q].buf.get();
  }
  {
    Timed tm(ctx, "k_gauss_xy", fields * 24.0 * double(n));
    if (!launch_gauss_xy_set(xy, d, t, s)) invalid("smoothing", "the batched x/y pass does not fit the grid");
  }
  {
    Timed tm(ctx, "k_gauss_z", fields * (24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0)));
    if (!launch_gauss_z_set(z, d, t, float(w1), float(w2), s)) invalid("smoothing", "the batched z pass does not fit the grid");
  }
  return out;
}

// The reports of the items of one batched inversion stage (coherent host memory).
struct InvStage {
  volatile InvertReport* reports = nullptr;
  size_t count = 0;
};

// invert_deferred for several (field, initial) pairs on one grid: each initial buffer becomes its
// inverse estimate (psi is updated in place), as invert_deferred returns it.
InvStage invert_items(HipBackend& b, const std::vector<const HField*>& fields, std::vector<HField*>& psi,
                      const InvertParams& p, InvertReport* report_dev, volatile InvertReport* report_host,
                      std::vector<DevBuf>& keep) {
  DeviceContext& ctx = b.ctx();
  const Grid& g = fields[0]->grid;
  const Dims d = dims_of(g);
  if (d.n() >= (1ll << 31)) invalid("invert", "grids of 2^31 voxels or more are not supported");
  float l[9];
  put(l, g.index_from_offset);
  const float3 isp = inv_spacing(g);
  const int nb = inv_blocks(d, int(ctx.tuning().get("inv.max_blocks", uint64_t(d.n()), 2048)));
  const size_t m = fields.size();
  InvItems it{};
  it.count = int(m);
  // per item: the residual field, the state and the two partial buffers
  DevBuf state(&ctx, m * (sizeof(InvertState) / sizeof(float)));
  DevBuf partials(&ctx, m * 4 * size_t(nb));
  for (size_t q = 0; q < m; ++q) {
    keep.emplace_back(&ctx, 3 * g.num_voxels());
    it.field[q] = fields[q]->buf.get();
    it.psi[q] = psi[q]->buf.get();
    it.res[q] = keep.back().get();
    it.state[q] = reinterpret_cast<InvertState*>(state.get()) + q;
    it.report[q] = report_dev + q;
    it.partials[q] = partials.get() + q * 4 * size_t(nb);
  }
  hipStream_t s = ctx.stream();
  const int max_it = int(p.max_iterations);
  volatile uint32_t* mail = ctx.mailbox_host();
  uint32_t* mail_dev = ctx.mailbox_device();
  const uint32_t tag = ctx.next_tag();
  {
    Timed t(ctx, "k_inv_begin", 36.0 * double(d.n()) * double(m));
    launch_inv_begin_items(it, d, l, isp, nb, max_it, s);
  }
  const auto all_stopped = [&] {
    for (size_t q = 0; q < m; ++q)
      if (mail[kInvItemWords + q] != tag) return false;
    return true;
  };
  const auto started = [&]() -> int {
    const uint32_t w = mail[1];
    return (w >> 8) == (tag & 0xffffffu) ? int(w & 0xffu) : 0;
  };
  // as invert_deferred: at most one launch queued beyond the one the device has started
  constexpr int kAhead = 1;
  for (int step = 0; step < max_it; ++step) {
    if (step > 0 && all_stopped()) break;
    if (step >= kAhead) {
      const auto t0 = clock_type::now();
      for (unsigned spins = 1; started() < step - kAhead + 1 && !all_stopped(); ++spins)
        if ((spins & 255) == 0 && clock_type::now() - t0 > std::chrono::milliseconds(50)) break;
      if (all_stopped()) break;
    }
    Timed t(ctx, "k_inv_iter", 0.0);
    launch_inv_iter_items(it, d, l, isp, nb, step, max_it, p.max_error_tolerance, p.mean_error_tolerance,
                          p.enforce_boundary_condition ? 1 : 0, mail_dev, tag, s);
  }
  keep.push_back(std::move(state));
  keep.push_back(std::move(partials));
  return InvStage{report_host, m};
}

InvertStats stats_of(const volatile InvertReport& r) {
  InvertStats st;
  st.iterations = uint32_t(r.executed);
  st.mean_error_norm = r.executed ? double(r.mean_error) : 0.0;
  st.max_error_norm = r.executed ? double(r.max_error) : 0.0;
  return st;
}

// The buffers of one level's iterations, sized for the `k` subjects active when the level starts
// and kept for the whole level (the q-th active subject uses the q-th slots, so the smaller active
// sets of later iterations fit), so that an iteration's launches can be captured into a HIP graph
// once and replayed: every pointer they take stays fixed. The graphs are keyed by the active set
// and the buffers its launches use (the compositions' buffers alternate between two per side).
struct LevelBuffers {
  size_t k = 0, nc = 0;
  Grid grid;
  std::vector<std::vector<HImage>> fmid, mmid;  // [k][nc]
  std::vector<std::vector<HField>> dfix, dmov;  // [k][nc]
  std::vector<HField> grads;                    // [2k] the channels combined (several channels)
  std::vector<HField> smoothed;                 // [2k]
  std::vector<DevBuf> tmp;                      // [2k] the x/y pass of the two-pass smoothing
  DevBuf norm_parts, coef, step_parts, factors;
  int nb = 0, tile_z = 0;
  size_t blocks = 0;
  char* cc_host = nullptr;
  float* pcc_dev = nullptr;
  int* pcount_dev = nullptr;
  char* small_host = nullptr;
  char* small_dev = nullptr;
  std::vector<std::pair<std::vector<const void*>, hipGraphExec_t>> graphs;

  LevelBuffers() = default;
  LevelBuffers(const LevelBuffers&) = delete;
  LevelBuffers& operator=(const LevelBuffers&) = delete;
  ~LevelBuffers() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  }
};

// Offsets in the iteration's coherent host words: per subject 2 maxima and a bad flag, then the
// reports of the two inversion stages (4 per subject).
size_t report_offset(size_t k) { return ((2 * k * sizeof(float) + k * sizeof(uint32_t)) + 15) / 16 * 16; }

std::unique_ptr<LevelBuffers> make_level_buffers(HipBackend& b, const SynRecipe& recipe, size_t k, size_t nc, const Grid& g) {
  DeviceContext& ctx = b.ctx();
  auto L = std::make_unique<LevelBuffers>();
  L->k = k;
  L->nc = nc;
  L->grid = g;
  const size_t n = g.num_voxels();
  const Dims d = dims_of(g);
  L->fmid.resize(k);
  L->mmid.resize(k);
  L->dfix.resize(k);
  L->dmov.resize(k);
  for (size_t q = 0; q < k; ++q)
    for (size_t c = 0; c < nc; ++c) {
      L->fmid[q].push_back(HImage{g, DevBuf(&ctx, n), 0.0f});
      L->mmid[q].push_back(HImage{g, DevBuf(&ctx, n), 0.0f});
      L->dfix[q].push_back(HField{g, DevBuf(&ctx, 3 * n)});
      L->dmov[q].push_back(HField{g, DevBuf(&ctx, 3 * n)});
    }
  for (size_t q = 0; q < 2 * k; ++q) {
    if (nc > 1) L->grads.push_back(HField{g, DevBuf(&ctx, 3 * n)});
    if (recipe.update_field_variance > 0.0) {
      L->smoothed.push_back(HField{g, DevBuf(&ctx, 3 * n)});
      L->tmp.emplace_back(&ctx, 3 * n);
    }
  }
  L->nb = norm_blocks(long(n));
  if (nc > 1) {
    L->norm_parts = DevBuf(&ctx, k * nc * 4 * size_t(L->nb));
    L->coef = DevBuf(&ctx, 2 * k * (kBatchChannels + 2));
  }
  L->step_parts = DevBuf(&ctx, k * 4 * size_t(L->nb));
  L->factors = DevBuf(&ctx, 2 * k);
  L->tile_z = cc_tile_for(d, int(ctx.tuning().get("cc.tile_z", n, b.cc_tile_z)), int(ctx.tuning().get("cc.min_blocks", n, 512)));
  const dim3 gr = cc_grid(d, L->tile_z);
  L->blocks = size_t(gr.x) * gr.y * gr.z;
  // coherent host memory for the whole level (later, smaller active sets fit in it)
  void* cc_dev = nullptr;
  L->cc_host = static_cast<char*>(ctx.mapped(6, k * nc * L->blocks * (sizeof(float) + sizeof(int)), &cc_dev));
  L->pcc_dev = static_cast<float*>(cc_dev);
  L->pcount_dev = reinterpret_cast<int*>(L->pcc_dev + k * nc * L->blocks);
  void* small_dev = nullptr;
  L->small_host = static_cast<char*>(ctx.mapped(7, report_offset(k) + 4 * k * sizeof(InvertReport) + 64, &small_dev));
  L->small_dev = static_cast<char*>(small_dev);
  return L;
}

// The taps of the update smoothing on grid g (as smooth_field_pair takes them) and whether the
// batched kernels take m fields of that grid at once.
bool batch_taps(DeviceContext& ctx, const Grid& g, size_t m, double variance, GaussTaps3& t) {
  int* radii[3] = {&t.rx, &t.ry, &t.rz};
  float* dst[3] = {t.x, t.y, t.z};
  for (int q = 0; q < 3; ++q) {
    const std::vector<double>& taps = ctx.host_taps(variance, g.size[q]);
    if (taps.size() > 17) return false;
    *radii[q] = int(taps.size() / 2);
    for (size_t k = 0; k < taps.size(); ++k) dst[q][k] = float(taps[k]);
  }
  const Dims d = dims_of(g);
  return gauss3_fits(d, t) && m <= size_t(kBatchFields) && 3ull * m * size_t(d.nz) <= 65535;
}

// smooth_fields into the level's buffers: the m updates `in` smoothed into L.smoothed with the
// same taps, pass shape and kernels (so the same bits), allocating nothing. The caller has checked
// batch_taps.
void smooth_into(DeviceContext& ctx, LevelBuffers& L, const std::vector<const float*>& in, double variance) {
  const Grid& g = L.grid;
  const Dims d = dims_of(g);
  const size_t n = g.num_voxels(), m = in.size();
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
  GaussTaps3 t{};
  if (!batch_taps(ctx, g, m, variance, t)) invalid("smoothing", "the batched smoothing does not fit the grid");
  hipStream_t s = ctx.stream();
  const double fields = double(m);
  FieldPtrs p{};
  p.count = int(m);
  for (size_t q = 0; q < m; ++q) {
    p.in[q] = in[q];
    p.orig[q] = in[q];
    p.out[q] = L.smoothed[q].buf.get();
  }
  if (t.rx == t.ry && t.ry == t.rz && ctx.tuning().get("gauss.one_pass", uint64_t(d.n()), gauss3_pays(d) ? 1 : 0)) {
    Timed tm(ctx, "k_gauss3", fields * (24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0)));
    if (launch_gauss3_set(p, d, t, float(w1), float(w2), s)) return;
  }
  FieldPtrs xy{}, z{};
  xy.count = z.count = int(m);
  for (size_t q = 0; q < m; ++q) {
    xy.in[q] = in[q];
    xy.out[q] = L.tmp[q].get();
    z.in[q] = L.tmp[q].get();
    z.orig[q] = in[q];
    z.out[q] = L.smoothed[q].buf.get();
  }
  {
    Timed tm(ctx, "k_gauss_xy", fields * 24.0 * double(n));
    if (!launch_gauss_xy_set(xy, d, t, s)) invalid("smoothing", "the batched x/y pass does not fit the grid");
  }
  Timed tm(ctx, "k_gauss_z", fields * (24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0)));
  if (!launch_gauss_z_set(z, d, t, float(w1), float(w2), s)) invalid("smoothing", "the batched z pass does not fit the grid");
}

// Whether the updates of m fields on grid g can be smoothed by smooth_into.
bool smooth_fits(DeviceContext& ctx, const Grid& g, size_t m, double variance) {
  GaussTaps3 t{};
  return variance <= 0.0 || batch_taps(ctx, g, m, variance, t);
}

// The launches of one iteration before the inversions: warps, CC pairs, channel weighting, update
// smoothing, step factors, compositions into the subjects' alternate buffers. Where the update
// smoothing fits the batched kernels they read and write only the level buffers, the subjects'
// fields and the level's coherent host words, and allocate nothing, so they can be captured.
void issue_updates(HipBackend& b, const SynRecipe& recipe, const std::vector<double>& normalised,
                   const std::vector<AffineTransform>& initial,
                   const std::vector<std::vector<SmoothCache<HipBackend>::Pair>>& smooth, std::vector<Subject>& subj,
                   const std::vector<size_t>& act, LevelBuffers& L) {
  DeviceContext& ctx = b.ctx();
  hipStream_t st = ctx.stream();
  const Grid& g = L.grid;
  const size_t k = act.size(), nc = normalised.size();
  const Dims d = dims_of(g);
  const size_t n = g.num_voxels();
  const double dn = double(n);
  // 1. every channel of every subject resampled into the middle (syn.rs:895-911)
  for (size_t c = 0; c < nc; ++c) {
    WarpItems wf{}, wm{};
    wf.count = wm.count = int(k);
    double src_bytes = 0.0;
    for (size_t q = 0; q < k; ++q) {
      const Subject& u = subj[act[q]];
      const HImage& F = *smooth[act[q]][c].fixed;
      const HImage& M = *smooth[act[q]][c].moving;
      L.fmid[q][c].bound = F.bound;
      L.mmid[q][c].bound = M.bound;
      wf.a[q] = warp_args(F.grid, *u.f2m_inv, nullptr, nullptr);
      wf.a[q].src = F.buf.get();
      wf.a[q].out = L.fmid[q][c].buf.get();
      wm.a[q] = warp_args(M.grid, *u.m2m_inv, nullptr, &initial[act[q]]);
      wm.a[q].src = M.buf.get();
      wm.a[q].out = L.mmid[q][c].buf.get();
      src_bytes += 4.0 * (std::min(dn, double(F.grid.num_voxels())) + std::min(dn, double(M.grid.num_voxels())));
    }
    Timed t(ctx, "k_warp", 32.0 * dn * double(k) + src_bytes);
    launch_warp_items(wf, st);
    launch_warp_items(wm, st);
  }
  // 2. the CC pair of every channel, both directions (syn.rs:337-420); block partials straight
  // into coherent host memory
  for (size_t c = 0; c < nc; ++c) {
    CcItems ci{};
    ci.count = int(k);
    for (size_t q = 0; q < k; ++q) {
      CcArgs& a = ci.a[q];
      a.fixed = L.fmid[q][c].buf.get();
      a.moving = L.mmid[q][c].buf.get();
      a.valid = nullptr;
      a.d = d;
      a.radius = int(recipe.cc_radius);
      a.tile_z = L.tile_z;
      a.scale_f = cc_split_scale(L.fmid[q][c].bound);
      a.scale_m = cc_split_scale(L.mmid[q][c].bound);
      a.step_f = 1.0f / a.scale_f;
      a.step_m = 1.0f / a.scale_m;
      a.bounds = nullptr;
      a.probes = probes_of(g);
      a.d_fixed = L.dfix[q][c].buf.get();
      a.d_moving = L.dmov[q][c].buf.get();
      a.partial_cc = L.pcc_dev + (q * nc + c) * L.blocks;
      a.partial_count = L.pcount_dev + (q * nc + c) * L.blocks;
    }
    Timed t(ctx, "k_cc2", 32.0 * dn * double(k));
    launch_cc_items(ci, st);
  }
  // the update of each side: the single channel's derivative, or the channels weighted on the
  // device as the host weighting does (syn.rs:203-297)
  std::vector<const float*> upd(2 * k);
  uint32_t* bad_dev = reinterpret_cast<uint32_t*>(L.small_dev + 2 * k * sizeof(float));
  if (nc == 1) {
    for (size_t q = 0; q < k; ++q) {
      upd[2 * q] = L.dfix[q][0].buf.get();
      upd[2 * q + 1] = L.dmov[q][0].buf.get();
    }
  } else {
    const int nb = L.nb;
    const float3 isp = inv_spacing(g);
    for (size_t c = 0; c < nc; ++c) {
      NormItems it{};
      it.count = int(k);
      for (size_t q = 0; q < k; ++q) {
        it.a[q] = L.dfix[q][c].buf.get();
        it.b[q] = L.dmov[q][c].buf.get();
        it.partial_max[q] = L.norm_parts.get() + (q * nc + c) * 4 * size_t(nb);
        it.partial_sq[q] = it.partial_max[q] + 2 * nb;
      }
      Timed t(ctx, "k_field_norms", 24.0 * dn * double(k));
      launch_field_norms_items(it, long(n), isp, nb, st);
    }
    WeightItems w{};
    w.channels = int(nc);
    w.count = int(k);
    for (size_t c = 0; c < nc; ++c) w.weight[c] = normalised[c];
    for (size_t q = 0; q < k; ++q)
      for (size_t c = 0; c < nc; ++c) w.partial_sq[q][c] = L.norm_parts.get() + (q * nc + c) * 4 * size_t(nb) + 2 * nb;
    launch_channel_weights(w, nb, L.coef.get(), bad_dev, st);
    CombineItems ci{};
    ci.channels = int(nc);
    ci.count = int(2 * k);
    for (size_t q = 0; q < k; ++q)
      for (int side = 0; side < 2; ++side) {
        for (size_t c = 0; c < nc; ++c) ci.d[2 * q + side][c] = (side == 0 ? L.dfix[q][c] : L.dmov[q][c]).buf.get();
        ci.out[2 * q + side] = L.grads[2 * q + side].buf.get();
        upd[2 * q + side] = L.grads[2 * q + side].buf.get();
      }
    Timed t(ctx, "k_combine", (12.0 * double(nc) + 12.0) * dn * double(2 * k));
    launch_combine_items(ci, L.coef.get(), long(3 * n), st);
  }
  // 3. smooth both updates of every subject: in the level's buffers, or (a grid the batched
  // kernels do not take; never captured) copies through the pair form
  std::vector<HField> held;
  if (recipe.update_field_variance > 0.0) {
    if (smooth_fits(ctx, g, 2 * k, recipe.update_field_variance)) {
      smooth_into(ctx, L, upd, recipe.update_field_variance);
      for (size_t q = 0; q < 2 * k; ++q) upd[q] = L.smoothed[q].buf.get();
    } else {
      std::vector<HField> copies;
      for (size_t q = 0; q < 2 * k; ++q) {
        copies.push_back(HField{g, DevBuf(&ctx, 3 * n)});
        check(hipMemcpyAsync(copies.back().buf.get(), upd[q], 3 * n * sizeof(float), hipMemcpyDeviceToDevice, st),
              "copy update");
      }
      held = smooth_fields(b, std::move(copies), recipe.update_field_variance);
      for (size_t q = 0; q < 2 * k; ++q) upd[q] = held[q].buf.get();
    }
  }
  // 4. each update scaled so its largest voxel step is the gradient step, applied in the
  // composition (syn.rs:422-445); 5. composed into the subject's alternate buffers
  const int nb = L.nb;
  NormItems it{};
  it.count = int(k);
  for (size_t q = 0; q < k; ++q) {
    it.a[q] = upd[2 * q];
    it.b[q] = upd[2 * q + 1];
    it.partial_max[q] = L.step_parts.get() + q * 4 * size_t(nb);
    it.partial_sq[q] = it.partial_max[q] + 2 * nb;
  }
  {
    Timed t(ctx, "k_field_norms", 24.0 * dn * double(k));
    launch_field_norms_items(it, long(n), inv_spacing(g), nb, st);
  }
  launch_step_factors_items(it, nb, recipe.gradient_step, L.factors.get(), reinterpret_cast<float*>(L.small_dev), st);
  FieldPtrs p{};
  p.count = int(2 * k);
  for (size_t q = 0; q < k; ++q) {
    Subject& u = subj[act[q]];
    for (int side = 0; side < 2; ++side) {
      p.in[2 * q + side] = upd[2 * q + side];
      p.orig[2 * q + side] = (side == 0 ? *u.f2m : *u.m2m).buf.get();
      p.out[2 * q + side] = (side == 0 ? *u.alt_f : *u.alt_m).buf.get();
    }
  }
  float l[9];
  put(l, g.index_from_offset);
  Timed t(ctx, "k_compose", 36.0 * dn * double(2 * k));
  launch_compose_set(p, d, l, L.factors.get(), st);
}

// One lockstep iteration of the subjects `act` on the level buffers L. Fills its[q] for each.
void iterate(HipBackend& b, const SynRecipe& recipe, const std::vector<double>& normalised,
             const std::vector<AffineTransform>& initial,
             const std::vector<std::vector<SmoothCache<HipBackend>::Pair>>& smooth, std::vector<Subject>& subj,
             const std::vector<size_t>& act, LevelBuffers& L, std::vector<SynIteration>& its) {
  DeviceContext& ctx = b.ctx();
  hipStream_t st = ctx.stream();
  const Grid& g = L.grid;
  const size_t k = act.size(), nc = normalised.size();
  const size_t n = g.num_voxels();
  const double dn = double(n);
  auto* bad_host = reinterpret_cast<uint32_t*>(L.small_host + 2 * k * sizeof(float));
  for (size_t q = 0; q < k; ++q) bad_host[q] = 0;
  // The updates: issued launch by launch, or (tuning knob batch.graphs = 1) replayed from a HIP
  // graph captured for this active set and these buffers. Replays give the same bits; on gfx950
  // with ROCm 7 they measured no faster at any level of the 1 mm pyramid (the host queues the
  // launches well ahead of the device), so direct launches are the default. Never while profiling
  // (the timing events belong to single launches) or when the total field is smoothed (that
  // smoothing allocates).
  const bool graphs = !ctx.profiling() && recipe.total_field_variance <= 0.0 &&
                      ctx.tuning().get("batch.graphs", uint64_t(n), 0) != 0 &&
                      smooth_fits(ctx, g, 2 * k, recipe.update_field_variance);
  if (graphs) {
    std::vector<const void*> key;
    for (size_t q : act) {
      const Subject& u = subj[q];
      key.push_back(reinterpret_cast<const void*>(q));
      for (const std::optional<HField>* f : {&u.f2m, &u.m2m, &u.f2m_inv, &u.m2m_inv, &u.alt_f, &u.alt_m})
        key.push_back((*f)->buf.get());
      for (const auto& pair : smooth[q]) {
        key.push_back(pair.fixed->buf.get());
        key.push_back(pair.moving->buf.get());
      }
    }
    hipGraphExec_t exec = nullptr;
    for (auto& e : L.graphs)
      if (e.first == key) exec = e.second;
    if (!exec) {
      hipGraph_t graph = nullptr;
      check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        issue_updates(b, recipe, normalised, initial, smooth, subj, act, L);
      } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      check(e, "hipGraphInstantiate");
      L.graphs.emplace_back(std::move(key), exec);
    }
    check(hipGraphLaunch(exec, st), "hipGraphLaunch");
  } else {
    issue_updates(b, recipe, normalised, initial, smooth, subj, act, L);
  }
  if (recipe.total_field_variance > 0.0)
    for (size_t q = 0; q < k; ++q) {
      Subject& u = subj[act[q]];
      u.alt_f = b.smooth_field(std::move(*u.alt_f), recipe.total_field_variance);
      u.alt_m = b.smooth_field(std::move(*u.alt_m), recipe.total_field_variance);
    }
  // 6. both inverses re-estimated: the new inverse from the old one, then the new field from it
  // (the composition, in its buffer)
  auto* reports_host = reinterpret_cast<volatile InvertReport*>(L.small_host + report_offset(L.k));
  auto* reports_dev = reinterpret_cast<InvertReport*>(L.small_dev + report_offset(L.k));
  const InvertParams inv_params;  // InvertParams::SYN
  std::vector<DevBuf> keep;
  std::vector<const HField*> fa, fb;
  std::vector<HField*> pa, pb;
  for (size_t q = 0; q < k; ++q) {
    Subject& u = subj[act[q]];
    fa.push_back(&*u.alt_f);
    pa.push_back(&*u.f2m_inv);
    fa.push_back(&*u.alt_m);
    pa.push_back(&*u.m2m_inv);
    fb.push_back(&*u.f2m_inv);
    pb.push_back(&*u.alt_f);
    fb.push_back(&*u.m2m_inv);
    pb.push_back(&*u.alt_m);
  }
  const InvStage sa = invert_items(b, fa, pa, inv_params, reports_dev, reports_host, keep);
  const InvStage sb = invert_items(b, fb, pb, inv_params, reports_dev + 2 * k, reports_host + 2 * k, keep);
  b.settle();
  for (size_t q = 0; q < k; ++q) {
    Subject& u = subj[act[q]];
    std::swap(u.f2m, u.alt_f);
    std::swap(u.m2m, u.alt_m);
  }
  const volatile float* maxima_host = reinterpret_cast<const volatile float*>(L.small_host);
  const bool profiling = ctx.profiling();
  for (size_t q = 0; q < k; ++q) {
    SynIteration& it_q = its[q];
    const double fmax = maxima_host[2 * q], mmax = maxima_host[2 * q + 1];
    if (!(std::isfinite(fmax) && std::isfinite(mmax))) invalid("syn", "the metric gradient is not finite");
    // stats in the order of syn_fields: F inverse, F forward, M inverse, M forward
    const volatile InvertReport* r[4] = {&sa.reports[2 * q], &sb.reports[2 * q], &sa.reports[2 * q + 1],
                                         &sb.reports[2 * q + 1]};
    for (int v = 0; v < 4; ++v) {
      it_q.inversions[v] = stats_of(*r[v]);
      if (profiling && r[v]->executed > 0) ctx.add_bytes("k_inv_iter", dn * (60.0 * double(r[v]->executed - 1) + 36.0));
      if (!(std::isfinite(it_q.inversions[v].mean_error_norm) && std::isfinite(it_q.inversions[v].max_error_norm)))
        invalid("syn", "the displacement field is not finite");
    }
    if (nc > 1 && bad_host[q]) invalid("syn", "the metric gradient is not finite");
    // the channels' values, each from its block partials in block order
    const float* pcc = reinterpret_cast<const float*>(L.cc_host);
    const int* pcount = reinterpret_cast<const int*>(L.cc_host + L.k * nc * L.blocks * sizeof(float));
    double weighted[2] = {0.0, 0.0};
    double first = 0.0;
    for (size_t c = 0; c < nc; ++c) {
      double total = 0.0;
      long long count = 0;
      for (size_t bl = 0; bl < L.blocks; ++bl) {
        total += double(pcc[(q * nc + c) * L.blocks + bl]);
        count += pcount[(q * nc + c) * L.blocks + bl];
      }
      if (count == 0) invalid("cc metric", "no virtual voxel maps inside both image buffers");
      const double v = -total / double(count);
      if (c == 0) first = v;
      weighted[0] += nc == 1 ? v : v * normalised[c];
      weighted[1] += nc == 1 ? v : v * normalised[c];
    }
    double scales[2] = {recipe.gradient_step, recipe.gradient_step};
    if (fmax > 0.0) scales[0] /= fmax;
    if (mmax > 0.0) scales[1] /= mmax;
    it_q.energy = 0.5 * (first + first);
    it_q.energy_weighted = 0.5 * (weighted[0] + weighted[1]);
    it_q.fixed_update_scale = scales[0];
    it_q.moving_update_scale = scales[1];
  }
}

}  // namespace

HipBackend::InvertedPair HipBackend::invert_pair_deferred(const HField& f0, HField initial0, const HField& f1,
                                                          HField initial1, const InvertParams& p) {
  if (!(f0.grid == f1.grid) || !(initial0.grid == f0.grid) || !(initial1.grid == f1.grid)) {
    auto a = invert_deferred(f0, std::move(initial0), p);
    auto b = invert_deferred(f1, std::move(initial1), p);
    return {std::move(a.first), std::move(b.first), std::move(a.second), std::move(b.second)};
  }
  // two adjacent slots of the inversion reports' ring (read when the statistics are asked for)
  constexpr size_t kReports = 64;
  void* report_dev = nullptr;
  auto* reports = static_cast<volatile InvertReport*>(ctx_->mapped(2, kReports * sizeof(InvertReport), &report_dev));
  const size_t slot = 2 * (size_t(ctx_->next_tag()) % (kReports / 2));
  std::vector<const HField*> fields{&f0, &f1};
  std::vector<HField*> psi{&initial0, &initial1};
  std::vector<DevBuf> keep;
  const InvStage st = invert_items(*this, fields, psi, p, static_cast<InvertReport*>(report_dev) + slot, reports + slot, keep);
  DeviceContext* ctx = ctx_.get();
  const double voxels = double(f0.grid.num_voxels());
  const bool profiling = ctx_->profiling();
  auto stats = [ctx, voxels, profiling](volatile InvertReport* r) {
    return Pending<InvertStats>(std::function<InvertStats()>([r, ctx, voxels, profiling] {
      if (profiling && r->executed > 0) ctx->add_bytes("k_inv_iter", voxels * (60.0 * double(r->executed - 1) + 36.0));
      return stats_of(*r);
    }));
  };
  return {std::move(initial0), std::move(initial1), stats(st.reports), stats(st.reports + 1)};
}

std::vector<SynFields> syn_fields_batch(HipBackend& b, const SynRecipe& recipe, const std::vector<AffineTransform>& initial,
                                        const std::vector<SmoothCache<HipBackend>*>& caches, const volatile int32_t* cancel,
                                        std::vector<std::optional<HField>>* forward,
                                        std::vector<std::optional<HField>>* inverse) {
  recipe.validate();
  const size_t ns = caches.size();
  if (ns == 0 || ns > size_t(kBatchItems))
    invalid("batch syn", "a batch holds 1 to " + std::to_string(kBatchItems) + " subjects");
  if (initial.size() != ns) invalid("batch syn", "one initial transform per subject is required");
  const size_t nc = caches[0]->channels();
  if (nc == 0) invalid("syn", "no channels");
  if (nc != recipe.channel_weights.size()) invalid("syn", "the number of channels differs from the number of channel weights");
  if (nc > size_t(kBatchChannels)) invalid("batch syn", "at most " + std::to_string(kBatchChannels) + " channels");
  for (size_t c = 0; c < nc; ++c)
    if (recipe.metric(c) != SynMetric::Cc) invalid("batch syn", "the batched loop takes CC channels only");
  const Grid fixed_grid = caches[0]->fixed_grid(0);
  for (size_t s = 0; s < ns; ++s) {
    if (!caches[s] || caches[s]->channels() != nc) invalid("batch syn", "every subject needs the same channels");
    for (size_t c = 0; c < nc; ++c)
      if (!caches[s]->fixed_grid(c).approx_eq(fixed_grid, 1e-9))
        invalid("batch syn", "the fixed images of the subjects differ in grid");
  }
  double wsum = 0.0;
  for (double w : recipe.channel_weights) wsum += w;
  std::vector<double> normalised;
  for (double w : recipe.channel_weights) normalised.push_back(w / wsum);
  std::vector<Subject> subj(ns);
  std::vector<SynFields> out(ns);
  for (size_t level = 0; level < recipe.levels(); ++level) {
    check_cancel(cancel);
    const auto t0 = clock_type::now();
    const auto factors = shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing);
    const Grid grid = level_virtual_grid(fixed_grid, factors);
    const double sigma = recipe.smoothing_sigmas[level];
    std::vector<std::vector<SmoothCache<HipBackend>::Pair>> smooth(ns);
    for (size_t s = 0; s < ns; ++s)
      for (size_t c = 0; c < nc; ++c) smooth[s].push_back(caches[s]->get(c, sigma, recipe.sigma_units));
    for (size_t s = 0; s < ns; ++s) {
      Subject& u = subj[s];
      if (!u.f2m) {
        u.f2m = b.zero_field(grid);
        u.f2m_inv = b.zero_field(grid);
        u.m2m = b.zero_field(grid);
        u.m2m_inv = b.zero_field(grid);
      } else {
        u.f2m = b.resample_field(*u.f2m, grid);
        u.f2m_inv = b.resample_field(*u.f2m_inv, grid);
        u.m2m = b.resample_field(*u.m2m, grid);
        u.m2m_inv = b.resample_field(*u.m2m_inv, grid);
      }
      u.monitor.emplace(recipe.convergence_window);
      u.trace = SynLevelTrace{};
      u.trace.level = level;
      u.trace.shrink_factors = factors;
      u.trace.grid_size = grid.size;
      u.current = 0;
      u.active = recipe.iterations[level] > 0;
      if (u.active) {
        u.alt_f = HField{grid, DevBuf(&b.ctx(), 3 * grid.num_voxels())};
        u.alt_m = HField{grid, DevBuf(&b.ctx(), 3 * grid.num_voxels())};
      }
    }
    std::unique_ptr<LevelBuffers> bufs;
    for (;;) {
      std::vector<size_t> act;
      for (size_t s = 0; s < ns; ++s)
        if (subj[s].active) act.push_back(s);
      if (act.empty()) break;
      check_cancel(cancel);
      const auto t_it = clock_type::now();
      std::vector<SynIteration> its(act.size());
      if (!bufs) bufs = make_level_buffers(b, recipe, act.size(), nc, grid);
      iterate(b, recipe, normalised, initial, smooth, subj, act, *bufs, its);
      const double secs = std::chrono::duration<double>(clock_type::now() - t_it).count();
      for (size_t q = 0; q < act.size(); ++q) {
        Subject& u = subj[act[q]];
        SynIteration& it = its[q];
        ++u.current;
        u.monitor->add_energy_value(it.energy);
        it.convergence_value = u.monitor->convergence_value();
        if (it.convergence_value < recipe.convergence_threshold) u.trace.converged = true;
        it.seconds = secs;
        u.trace.iterations.push_back(it);
        if (u.trace.converged || u.current >= recipe.iterations[level]) u.active = false;
      }
    }
    bufs.reset();
    const double level_secs = std::chrono::duration<double>(clock_type::now() - t0).count();
    for (size_t s = 0; s < ns; ++s) {
      subj[s].alt_f.reset();
      subj[s].alt_m.reset();
      subj[s].trace.seconds = level_secs;
      out[s].levels.push_back(std::move(subj[s].trace));
    }
  }
  // syn.rs:969-980: SyN = Minv o F forward, Finv o M inverse
  for (size_t s = 0; s < ns; ++s) {
    Subject& u = subj[s];
    if (forward) (*forward)[s] = b.compose(*u.m2m_inv, *u.f2m);
    if (inverse) (*inverse)[s] = b.compose(*u.f2m_inv, *u.m2m);
  }
  return out;
}

std::vector<HipBackend*> WorkerPool::take(size_t count) {
  while (workers_.size() < count) workers_.push_back(std::make_unique<HipBackend>(device_));
  std::vector<HipBackend*> out;
  for (size_t w = 0; w < count; ++w) out.push_back(workers_[w].get());
  return out;
}

void run_linear_fits(const std::vector<HipBackend*>& workers, const RegisterPlan& plan, std::vector<LinearFitJob>& jobs,
                     const volatile int32_t* cancel) {
  if (workers.empty()) invalid("linear fits", "no workers");
  std::atomic<size_t> next{0};
  std::atomic<bool> failed{false};
  std::mutex error_mutex;
  std::exception_ptr error;
  auto work = [&](HipBackend& b) {
    try {
      check(hipSetDevice(b.ctx().device()), "hipSetDevice");
      for (size_t i = next++; i < jobs.size() && !failed; i = next++) {
        LinearFitJob& job = jobs[i];
        auto t = clock_type::now();
        auto fixed_pre = job.fixed_pre ? job.fixed_pre
                                       : std::make_shared<const HImage>(
                                             b.preprocess_image(*job.fixed_raw, plan.winsorize_lower, plan.winsorize_upper));
        auto moving_pre = job.moving_pre ? job.moving_pre
                                         : std::make_shared<const HImage>(
                                               b.preprocess_image(*job.moving_raw, plan.winsorize_lower, plan.winsorize_upper));
        job.seconds_preprocess = std::chrono::duration<double>(clock_type::now() - t).count();
        t = clock_type::now();
        SmoothCache<HipBackend> cache(b, std::vector<SmoothCache<HipBackend>::Pair>{{fixed_pre, moving_pre}});
        if (job.shared) cache.share_fixed(job.shared);
        AffineTransform start = plan.initial_affine;
        if (plan.initial_kind == InitialKind::CentreOfMass) {
          const Vec3 fc = job.fixed_centre ? *job.fixed_centre : b.centre_of_gravity(*job.fixed_raw);
          const Vec3 mc = b.centre_of_gravity(*job.moving_raw);
          start = AffineTransform{kIdentity3, vsub(mc, fc), fc};
        } else if (plan.initial_kind == InitialKind::Identity) {
          start = AffineTransform{};
        }
        if (plan.run_linear) {
          HipLinearEngine engine(b);
          job.out = run_linear_chain(engine, plan.rigid, plan.affine, start, cache, cancel);
        } else {
          job.out.initial = job.out.total = start;
        }
        b.ctx().sync();
        job.seconds_linear = std::chrono::duration<double>(clock_type::now() - t).count();
      }
      b.ctx().sync();
    } catch (...) {
      failed = true;
      std::lock_guard<std::mutex> lock(error_mutex);
      if (!error) error = std::current_exception();
      try {
        b.ctx().sync();
      } catch (...) {
      }
    }
  };
  const size_t nw = std::min(workers.size(), jobs.size());
  std::vector<std::thread> threads;
  for (size_t w = 1; w < nw; ++w) threads.emplace_back(work, std::ref(*workers[w]));
  if (nw > 0) work(*workers[0]);
  for (auto& th : threads) th.join();
  if (error) std::rethrow_exception(error);
}

std::vector<RegisterOutput> register_batch(HipBackend& b, const std::vector<HipBackend*>& workers, const RegisterPlan& plan,
                                           const std::vector<const HImage*>& fixed,
                                           const std::vector<std::vector<const HImage*>>& moving,
                                           const volatile int32_t* cancel, std::vector<std::optional<HField>>* forward,
                                           std::vector<std::optional<HField>>* inverse) {
  auto secs = [](clock_type::time_point a) { return std::chrono::duration<double>(clock_type::now() - a).count(); };
  const size_t ns = moving.size(), nc = fixed.size();
  if (ns == 0) invalid("register batch", "no subjects");
  if (nc == 0) invalid("register", "fixed and moving must have the same, non-zero number of channels");
  for (size_t s = 0; s < ns; ++s)
    if (moving[s].size() != nc) invalid("register", "fixed and moving must have the same, non-zero number of channels");
  if (plan.run_syn && plan.syn.channel_weights.size() != nc)
    invalid("stage plan", "the number of channels differs from the number of channel weights");
  for (size_t c = 0; c < nc; ++c) {
    if (!fixed[c]) invalid("register", "missing image");
    for (size_t s = 0; s < ns; ++s)
      if (!moving[s][c]) invalid("register", "missing image");
  }
  if (forward) {
    forward->clear();
    forward->resize(ns);
  }
  if (inverse) {
    inverse->clear();
    inverse->resize(ns);
  }
  std::vector<RegisterOutput> out(ns);
  // the template: preprocessed, its centre of mass taken and every level it is smoothed at built
  // once (in the order a single run asks for them: rigid, affine, then SyN), on this context
  auto t = clock_type::now();
  std::vector<std::shared_ptr<const HImage>> tmpl;
  for (size_t c = 0; c < nc; ++c)
    tmpl.push_back(std::make_shared<const HImage>(b.preprocess_image(*fixed[c], plan.winsorize_lower, plan.winsorize_upper)));
  std::vector<std::vector<std::shared_ptr<const HImage>>> subj(ns);
  for (size_t s = 0; s < ns; ++s)
    for (size_t c = 0; c < nc; ++c)
      subj[s].push_back(
          std::make_shared<const HImage>(b.preprocess_image(*moving[s][c], plan.winsorize_lower, plan.winsorize_upper)));
  const double pre_secs = secs(t);
  auto shared = std::make_shared<SmoothCache<HipBackend>::SharedFixed>();
  auto level = [&](size_t c, double sigma, SigmaUnits units) {
    if (sigma <= 0.0 || shared->count({c, sigma})) return;
    (*shared)[{c, sigma}] = std::make_shared<const HImage>(b.smooth_image(*tmpl[c], sigma, units));
  };
  if (plan.run_linear)
    for (const LinearRecipe* r : {&plan.rigid, &plan.affine})
      for (double sigma : r->smoothing_sigmas) level(0, sigma, r->sigma_units);
  if (plan.run_syn)
    for (double sigma : plan.syn.smoothing_sigmas)
      for (size_t c = 0; c < nc; ++c) level(c, sigma, plan.syn.sigma_units);
  std::optional<Vec3> fc;
  if (plan.initial_kind == InitialKind::CentreOfMass) fc = b.centre_of_gravity(*fixed[0]);
  b.ctx().sync();
  // the subjects' linear chains, side by side on the workers
  std::vector<LinearFitJob> jobs(ns);
  for (size_t s = 0; s < ns; ++s) {
    jobs[s].fixed_raw = fixed[0];
    jobs[s].moving_raw = moving[s][0];
    jobs[s].fixed_pre = tmpl[0];
    jobs[s].moving_pre = subj[s][0];
    jobs[s].shared = shared;
    jobs[s].fixed_centre = fc;
  }
  run_linear_fits(workers, plan, jobs, cancel);
  for (size_t s = 0; s < ns; ++s) {
    RegisterOutput& o = out[s];
    o.seconds_preprocess = pre_secs;
    o.linear = jobs[s].out;
    o.initial = jobs[s].out.initial;
    o.total = jobs[s].out.total;
    o.seconds_linear = jobs[s].seconds_linear;
  }
  if (plan.run_syn) {
    b.release_idle_buffers();
    std::vector<std::unique_ptr<SmoothCache<HipBackend>>> caches;
    for (size_t s = 0; s < ns; ++s) {
      std::vector<SmoothCache<HipBackend>::Pair> resident;
      for (size_t c = 0; c < nc; ++c) resident.push_back({tmpl[c], subj[s][c]});
      caches.push_back(std::make_unique<SmoothCache<HipBackend>>(b, std::move(resident)));
      caches.back()->share_fixed(shared);
    }
    for (size_t s0 = 0; s0 < ns; s0 += size_t(kBatchItems)) {
      const size_t s1 = std::min(ns, s0 + size_t(kBatchItems));
      std::vector<AffineTransform> init;
      std::vector<SmoothCache<HipBackend>*> group;
      for (size_t s = s0; s < s1; ++s) {
        init.push_back(out[s].total);
        group.push_back(caches[s].get());
      }
      std::vector<std::optional<HField>> fwd(s1 - s0), inv(s1 - s0);
      t = clock_type::now();
      std::vector<SynFields> syn = syn_fields_batch(b, plan.syn, init, group, cancel, forward ? &fwd : nullptr,
                                                    inverse ? &inv : nullptr);
      const double batch_secs = secs(t);
      for (size_t s = s0; s < s1; ++s) {
        out[s].syn = std::move(syn[s - s0]);
        out[s].seconds_syn = batch_secs;
        if (forward) (*forward)[s] = std::move(fwd[s - s0]);
        if (inverse) (*inverse)[s] = std::move(inv[s - s0]);
      }
    }
  }
  for (auto& o : out) o.seconds_total = o.seconds_preprocess + o.seconds_linear + o.seconds_syn;
  return out;
}

}  // namespace fw::hip
