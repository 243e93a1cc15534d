// Host side of the Mattes MI kernels: the linear-stage level (the reference's GpuMiLevel,
// crates/fastwarp-gpu/src/wgpu_backend/mi.rs:144-524) and the dense SyN pair
// (wgpu_backend/mi/dense.rs:241-487). The device returns exact integer histograms; the host
// finishes them in f64 with the same code path for both (metric/mi.rs:702-769).
#include <atomic>
#include <cstring>

#include "backend.hpp"
#include "geometry.hpp"

namespace fw::hip {

namespace {

std::vector<double> joint_from(const unsigned long long* h, size_t bins) {
  std::vector<double> j(bins * bins);
  for (size_t i = 0; i < j.size(); ++i) j[i] = double(h[i]) / kMiFixedPoint;
  return j;
}

std::vector<double> marginal_from(const unsigned long long* h, size_t bins) {
  std::vector<double> m(bins);
  for (size_t i = 0; i < bins; ++i) m[i] = double(h[bins * bins + i]);
  return m;
}

}  // namespace

// --------------------------------------------------------------------- linear stage level ---

MiLevel::MiLevel(HipBackend& backend, const HImage& moving, const HostVec<Vec3>& points,
                 const HostVec<uint32_t>& fixed_bin, const MiBinning& binning)
    : ctx_(&backend.ctx()), moving_(&moving), count_(points.size()), binning_(binning) {
  if (binning.bins < 5 || binning.bins > size_t(kMiMaxBins))
    invalid("hip backend", "the linear Mattes metric supports 5 to 64 histogram bins, not " + std::to_string(binning.bins));
  if (fixed_bin.size() != points.size()) invalid("hip backend", "one fixed bin per sample point is required");
  if (points.size() > size_t(1) << 30) invalid("hip backend", "too many sample points");
  // the anchor: the points' mean, summed over fixed chunks combined in chunk order
  constexpr size_t kGrain = size_t(1) << 14;
  const size_t m = points.size();
  std::vector<Vec3> part((m + kGrain - 1) / kGrain, Vec3{0, 0, 0});
  parallel_chunks(m, kGrain, [&](size_t c, size_t b, size_t e) {
    Vec3 acc{0, 0, 0};
    for (size_t i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) acc[a] += points[i][a];
    part[c] = acc;
  });
  Vec3 mean{0, 0, 0};
  for (const Vec3& p : part)
    for (int a = 0; a < 3; ++a) mean[a] += p[a];
  if (m > 0)
    for (int a = 0; a < 3; ++a) mean[a] /= double(m);
  anchor_ = mean;
  const size_t floats = 4 * std::max<size_t>(m, 1);
  auto* rec = static_cast<float*>(ctx_->pinned(floats * sizeof(float)));
  std::atomic<bool> bad_bin{false};
  parallel_for(m, [&](size_t b, size_t e) {
    bool bad = false;
    for (size_t i = b; i < e; ++i) {
      for (int a = 0; a < 3; ++a) rec[4 * i + a] = float(points[i][a] - anchor_[a]);
      const uint32_t fb = fixed_bin[i];
      bad |= fb < 2 || fb > binning.bins - 3;
      const int32_t bin = int32_t(fb);
      std::memcpy(&rec[4 * i + 3], &bin, 4);
    }
    if (bad) bad_bin.store(true);
  });
  if (bad_bin.load()) invalid("hip backend", "fixed bin outside [2, bins - 3]");
  samples_ = DevBuf(ctx_, floats);
  check(hipMemcpyAsync(samples_.get(), rec, floats * sizeof(float), hipMemcpyHostToDevice, ctx_->stream()), "upload samples");
  ctx_->sync();
}

// The arguments of one histogram pass of transform t; returns the pass's block count.
int MiLevel::pass_args(const LinearMovingTransform& t, MiSampledArgs& a, size_t histograms) {
  const Grid& mg = moving_->grid;
  const AffineTransform T = t.collapsed();
  const Mat3 A = mat3_mul(mg.index_from_offset, T.matrix);
  const Vec3 b = mat3_mul_vec(mg.index_from_offset, vsub(vadd(mat3_mul_vec(T.matrix, anchor_), T.offset()), mg.origin));
  a = MiSampledArgs{};
  a.samples = reinterpret_cast<const float4*>(samples_.get());
  a.count = int(count_);
  a.moving = moving_->buf.get();
  a.md = dims_of(mg);
  put(a.a, A);
  put(a.b, b);
  a.inv_bin_size = float(1.0 / binning_.moving_bin_size);
  a.moving_min = float(binning_.moving_min);
  a.moving_max = float(binning_.moving_max);
  a.bins = int(binning_.bins);
  const size_t cells = binning_.bins * binning_.bins + binning_.bins + 1;
  // a batched pass bins every sample under each transform, so its blocks (and their slots) have a
  // cap of their own (tuning knob mi.batch_blocks; the single pass's cap by default)
  const uint64_t cap = ctx_->tuning().get("mi.max_blocks", count_, 512);
  const int blocks = mi_sampled_blocks(int(count_), int(histograms > 1 ? ctx_->tuning().get("mi.batch_blocks", count_, cap) : cap));
  a.threads = int(ctx_->tuning().get("mi.threads", count_, 0));
  a.hist_slots =
      static_cast<unsigned long long*>(ctx_->scratch(6, size_t(blocks) * histograms * cells * sizeof(unsigned long long)));
  return blocks;
}

// Adds the slots of `k` histograms into device memory and finishes them there (k_mi_finish)
// into `out` (a device view); with `signal` the finish then stores `tag` into done words [0, k).
void MiLevel::launch_finish(const unsigned long long* slots, int blocks, size_t k, float* ratio, MiFinishOut* out,
                            bool signal, uint32_t tag) {
  const size_t bins = binning_.bins;
  const size_t all = k * (bins * bins + bins + 1);
  auto* dev_hist = static_cast<unsigned long long*>(ctx_->scratch(8, all * sizeof(unsigned long long)));
  hipStream_t s = ctx_->stream();
  {
    Timed tm(*ctx_, "k_mi_sum_slots", 8.0 * double(all) * double(blocks + 1));
    launch_mi_sum_slots(slots, blocks, int(all), dev_hist, nullptr, 0, s);
  }
  Timed tm(*ctx_, "k_mi_finish", 8.0 * double(all) + (ratio ? 4.0 * double(bins * bins) : 0.0));
  launch_mi_finish(dev_hist, int(k), int(bins), binning_.moving_bin_size, out, ratio,
                   signal ? ctx_->done_words_device() : nullptr, tag, s);
}

// One evaluation of k transforms (k = 1: the pass arguments in `a`; else `batch`), finished into
// `out` (device view): the histogram pass, then the slot sum and the finish, in one launch when
// the pass has at most kMiFusedSlots slots (tuning knob mi.fused_slots) and kMiFusedBins bins.
void MiLevel::evaluate(const MiSampledArgs& a, const MiBatch& batch, size_t k, int blocks, float* ratio, MiFinishOut* out,
                       bool signal, uint32_t tag) {
  hipStream_t s = ctx_->stream();
  if (k == 1) {
    Timed tm(*ctx_, "k_mi_sampled_hist", 20.0 * double(count_));
    launch_mi_sampled_hist(a, blocks, s);
  } else {
    Timed tm(*ctx_, "k_mi_sampled_hist_batch", (16.0 + 4.0 * double(k)) * double(count_));
    launch_mi_sampled_hist_batch(a, batch, int(k), blocks, s);
  }
  const size_t bins = binning_.bins;
  if (blocks <= int(ctx_->tuning().get("mi.fused_slots", count_, kMiFusedSlots)) && bins <= size_t(kMiFusedBins)) {
    const double cells = double(k * (bins * bins + bins + 1));
    Timed tm(*ctx_, "k_mi_sum_finish", 8.0 * cells * double(blocks) + (ratio ? 4.0 * double(bins * bins) : 0.0));
    launch_mi_sum_finish(a.hist_slots, blocks, int(k), int(bins), binning_.moving_bin_size, out, ratio,
                         signal ? ctx_->done_words_device() : nullptr, tag, s);
    return;
  }
  launch_finish(a.hist_slots, blocks, k, ratio, out, signal, tag);
}

// The finish of `k` histograms in coherent host memory: host view, with the device view in *dev.
volatile MiFinishOut* MiLevel::finish_slots(size_t k, MiFinishOut** dev) {
  void* d = nullptr;
  auto* host = static_cast<volatile MiFinishOut*>(ctx_->mapped(5, k * sizeof(MiFinishOut), &d));
  *dev = static_cast<MiFinishOut*>(d);
  return host;
}

namespace {

double finished_value(const volatile MiFinishOut& o) {
  if (o.status == 1) invalid("mattes", "all samples map outside the moving image buffer");
  if (o.status == 2) invalid("mattes", "joint PDF summed to zero");
  return o.value;
}

}  // namespace

void MiLevel::histogram_cells(const LinearMovingTransform& t, std::vector<double>& joint, std::vector<double>& fixed_marginal,
                              uint64_t& valid) {
  MiSampledArgs a;
  const int blocks = pass_args(t, a, 1);
  const size_t bins = binning_.bins;
  const size_t cells = bins * bins + bins + 1;
  hipStream_t s = ctx_->stream();
  {
    Timed tm(*ctx_, "k_mi_sampled_hist", 20.0 * double(count_));
    launch_mi_sampled_hist(a, blocks, s);
  }
  // the sums go straight to coherent host memory; the completion words say when they are there
  void* dev_hist = nullptr;
  auto* host = static_cast<unsigned long long*>(ctx_->mapped(0, cells * sizeof(unsigned long long), &dev_hist));
  const uint32_t tag = ctx_->next_tag();
  {
    Timed tm(*ctx_, "k_mi_sum_slots", 8.0 * double(cells) * double(blocks + 1));
    launch_mi_sum_slots(a.hist_slots, blocks, int(cells), static_cast<unsigned long long*>(dev_hist),
                        ctx_->done_words_device(), tag, s);
  }
  ctx_->wait_done(mi_sum_slots_blocks(int(cells)), tag);
  valid = host[cells - 1];
  joint = joint_from(host, bins);
  fixed_marginal = marginal_from(host, bins);
}

double MiLevel::value(const LinearMovingTransform& t) {
  MiSampledArgs a;
  const int blocks = pass_args(t, a, 1);
  MiFinishOut* dev = nullptr;
  volatile MiFinishOut* out = finish_slots(1, &dev);
  const uint32_t tag = ctx_->next_tag();
  evaluate(a, MiBatch{}, 1, blocks, nullptr, dev, true, tag);
  ctx_->wait_done(1, tag);
  return finished_value(out[0]);
}

std::vector<double> MiLevel::values(const std::vector<LinearMovingTransform>& ts) {
  std::vector<double> out;
  out.reserve(ts.size());
  const size_t cap = size_t(mi_batch_capacity(int(binning_.bins)));
  for (size_t i = 0; i < ts.size();) {
    const size_t k = std::min(cap, ts.size() - i);
    if (k == 1) out.push_back(value(ts[i]));
    else batch_values(&ts[i], k, out);
    i += k;
  }
  return out;
}

// Up to kMiMaxBatch transforms in one histogram pass, one slot sum and one finish.
void MiLevel::batch_values(const LinearMovingTransform* ts, size_t k, std::vector<double>& out) {
  const Grid& mg = moving_->grid;
  MiSampledArgs a{};
  const int blocks = pass_args(ts[0], a, k);
  MiBatch batch{};
  for (size_t q = 0; q < k; ++q) {
    const AffineTransform T = ts[q].collapsed();
    const Mat3 A = mat3_mul(mg.index_from_offset, T.matrix);
    const Vec3 b = mat3_mul_vec(mg.index_from_offset, vsub(vadd(mat3_mul_vec(T.matrix, anchor_), T.offset()), mg.origin));
    put(batch.a[q], A);
    put(batch.b[q], b);
  }
  MiFinishOut* dev = nullptr;
  volatile MiFinishOut* fin = finish_slots(k, &dev);
  const uint32_t tag = ctx_->next_tag();
  evaluate(a, batch, k, blocks, nullptr, dev, true, tag);
  ctx_->wait_done(int(k), tag);
  for (size_t q = 0; q < k; ++q) out.push_back(finished_value(fin[q]));
}

std::pair<double, std::vector<double>> MiLevel::value_and_derivative(const LinearMovingTransform& t) {
  MiSampledArgs a;
  const int hist_blocks = pass_args(t, a, 1);
  const size_t bins = binning_.bins;
  hipStream_t s = ctx_->stream();
  // the finish leaves the ratio weights and its result on the device, where the derivative pass
  // reads the weights and its last kernel passes the result to the host with the derivative: one
  // wait for both
  auto* dev_ratio = static_cast<float*>(ctx_->scratch(1, bins * bins * sizeof(float)));
  auto* dev_fin = static_cast<MiFinishOut*>(ctx_->scratch(9, sizeof(MiFinishOut)));
  evaluate(a, MiBatch{}, 1, hist_blocks, dev_ratio, dev_fin, false, 0);
  a.ratio = dev_ratio;
  a.probes = probes_of(moving_->grid);
  put(a.outer, t.outer.matrix);
  put(a.rel_center, vsub(anchor_, t.optimised.center));
  const double* p = t.optimised.params.data();
  a.trig[0] = float(std::cos(p[0]));
  a.trig[1] = float(std::sin(p[0]));
  a.trig[2] = float(std::cos(p[1]));
  a.trig[3] = float(std::sin(p[1]));
  a.trig[4] = float(std::cos(p[2]));
  a.trig[5] = float(std::sin(p[2]));
  a.kind = t.optimised.kind == ParamKind::Affine ? 1 : 0;
  if (t.optimised.kind == ParamKind::Euler3d && t.optimised.zyx)
    invalid("hip backend", "the ZYX Euler order is not supported by the device derivative");
  const int blocks = mi_deriv_blocks(int(count_), int(ctx_->tuning().get("mi.deriv_blocks", count_, kMiDerivMaxBlocks)));
  const size_t np = t.optimised.count();
  a.partial = static_cast<double*>(ctx_->scratch(7, size_t(blocks) * 12 * sizeof(double)));
  {
    Timed tm(*ctx_, "k_mi_sampled_deriv", 20.0 * double(count_));
    launch_mi_sampled_deriv(a, blocks, s);
  }
  // the blocks' partials are added on the device, straight into coherent host memory; the
  // completion word says when the sums (and, before them, the finish) are there
  void* dev_sum = nullptr;
  auto* host = static_cast<double*>(ctx_->mapped(1, 12 * sizeof(double), &dev_sum));
  MiFinishOut* fin_dev = nullptr;
  volatile MiFinishOut* fin = finish_slots(1, &fin_dev);
  const uint32_t tag = ctx_->next_tag();
  {
    Timed tm(*ctx_, "k_mi_deriv_finish", 8.0 * 12.0 * double(blocks));
    launch_mi_deriv_finish(a.partial, blocks, int(np), static_cast<double*>(dev_sum), ctx_->done_words_device(), tag, s,
                           dev_fin, fin_dev);
  }
  ctx_->wait_done(1, tag);
  const double value = finished_value(fin[0]);
  std::vector<double> d(host, host + np);
  return {value, d};
}

// ---------------------------------------------------------------------------- dense pair ---

MetricPair<HField> HipBackend::mi_pair(const HImage& fixed, const HImage& moving, const HImage* fixed_mask,
                                       const HImage* moving_mask, size_t bins, bool want_norms) {
  if (fixed.grid != moving.grid) invalid("mattes", "the images are not on one grid");
  for (const HImage* m : {fixed_mask, moving_mask})
    if (m && m->grid != fixed.grid) invalid("mattes", "a mask is not on the grid of the images");
  if (bins < 5 || bins > 48) invalid("hip backend", "the dense Mattes metric supports 5 to 48 histogram bins, not " + std::to_string(bins));
  const Grid& g = fixed.grid;
  const long long n = (long long)g.num_voxels();
  hipStream_t s = ctx_->stream();
  MiDenseArgs a{};
  a.f = fixed.buf.get();
  a.m = moving.buf.get();
  a.fmask = fixed_mask ? fixed_mask->buf.get() : nullptr;
  a.mmask = moving_mask ? moving_mask->buf.get() : nullptr;
  a.d = dims_of(g);
  a.bins = int(bins);
  const int rb = mi_dense_blocks(n);
  a.range_partial = static_cast<float*>(ctx_->scratch(0, size_t(rb) * 4 * sizeof(float)));
  {
    Timed tm(*ctx_, "k_mi_dense_range", 8.0 * double(n));
    launch_mi_dense_range(a, rb, s);
  }
  auto* rp = static_cast<float*>(ctx_->pinned(size_t(rb) * 4 * sizeof(float)));
  check(hipMemcpyAsync(rp, a.range_partial, size_t(rb) * 4 * sizeof(float), hipMemcpyDeviceToHost, s), "read ranges");
  ctx_->sync();
  float fmin = INFINITY, fmax = -INFINITY, mmin = INFINITY, mmax = -INFINITY;
  for (int b = 0; b < rb; ++b) {
    fmin = std::min(fmin, rp[4 * b]);
    fmax = std::max(fmax, rp[4 * b + 1]);
    mmin = std::min(mmin, rp[4 * b + 2]);
    mmax = std::max(mmax, rp[4 * b + 3]);
  }
  if (!(fmin <= fmax))
    invalid("mattes", fixed_mask ? "no finite fixed-image voxel lies inside the fixed mask" : "the fixed image has no finite voxel");
  if (!(mmin <= mmax))
    invalid("mattes", moving_mask ? "no finite moving-image voxel lies inside the moving mask" : "the moving image has no finite voxel");
  const MiBinning b0 = MiBinning::from_ranges(bins, fmin, fmax, mmin, mmax);
  const MiBinning b1 = MiBinning::from_ranges(bins, mmin, mmax, fmin, fmax);
  const MiBinning* bn[2] = {&b0, &b1};
  for (int dir = 0; dir < 2; ++dir) {
    const std::vector<float> thr = bn[dir]->fixed_thresholds();
    for (size_t q = 0; q < thr.size(); ++q) a.fixed_thr[dir][q] = thr[q];
    a.moving_min[dir] = float(bn[dir]->moving_min);
    a.moving_max[dir] = float(bn[dir]->moving_max);
    a.inv_bin_size[dir] = float(1.0 / bn[dir]->moving_bin_size);
  }
  const size_t stride = bins * bins + bins + 1;
  a.hist = static_cast<unsigned long long*>(ctx_->scratch(1, 2 * stride * sizeof(unsigned long long)));
  check(hipMemsetAsync(a.hist, 0, 2 * stride * sizeof(unsigned long long), s), "clear histograms");
  {
    Timed tm(*ctx_, "k_mi_dense_hist", 8.0 * double(n));
    launch_mi_dense_hist(a, mi_dense_blocks(n), s);
  }
  auto* h = static_cast<unsigned long long*>(ctx_->pinned(2 * stride * sizeof(unsigned long long)));
  check(hipMemcpyAsync(h, a.hist, 2 * stride * sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "read histograms");
  ctx_->sync();
  MetricPair<HField> out;
  std::vector<float> ratio(2 * bins * bins);
  for (int dir = 0; dir < 2; ++dir) {
    const unsigned long long* hd = h + dir * stride;
    const uint64_t valid = hd[stride - 1];
    if (valid == 0 && (fixed_mask || moving_mask))
      invalid("mattes", "no voxel of the registration level lies inside both the fixed and the moving mask");
    HistogramFinish fin = finish_histogram(bins, bn[dir]->moving_bin_size, valid, joint_from(hd, bins), marginal_from(hd, bins), true);
    out.values[dir] = fin.value;
    for (size_t q = 0; q < bins * bins; ++q) ratio[dir * bins * bins + q] = float(fin.ratio[q]);
  }
  auto* dev_ratio = static_cast<float*>(ctx_->scratch(2, ratio.size() * sizeof(float)));
  check(hipMemcpyAsync(dev_ratio, ratio.data(), ratio.size() * sizeof(float), hipMemcpyHostToDevice, s), "upload ratio");
  a.ratio = dev_ratio;
  a.probes = probes_of(g);
  out.fields[0] = HField{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  out.fields[1] = HField{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  a.d0 = out.fields[0].buf.get();
  a.d1 = out.fields[1].buf.get();
  {
    // F and M (and the masks) read, both derivative fields written
    Timed tm(*ctx_, "k_mi_dense_deriv", (32.0 + (a.fmask ? 4.0 : 0.0) + (a.mmask ? 4.0 : 0.0)) * double(n));
    launch_mi_dense_deriv(a, s);
  }
  if (want_norms) {
    out.norms[0] = l2_norm(out.fields[0]);
    out.norms[1] = l2_norm(out.fields[1]);
  } else {
    ctx_->sync();  // the ratio upload reads host memory that goes out of scope
  }
  return out;
}

}  // namespace fw::hip
