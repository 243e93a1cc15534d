// HipBackend: the MI355X implementation of the reference's `Backend` trait
// (crates/fastwarp-reg/src/backend.rs:39-305) and `LinearEngine` (linear.rs:288-320).
//
// One DeviceContext is one HIP device with one stream and a size-keyed buffer pool. Every op is
// enqueued on that stream, so a buffer returned to the pool can be handed out again at once: later
// work on the stream is ordered after its earlier uses. Host round trips are limited to the
// scalars the drivers consume (metric values, norms, inversion statistics, MI histograms).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../host/backend_types.hpp"
#include "../host/tuning.hpp"
#include "../host/grid.hpp"
#include "../host/mi_host.hpp"
#include "../host/pyramid.hpp"
#include "../host/transform.hpp"
#include "kernels.hpp"
#include "mi_kernels.hpp"

namespace fw::hip {

void check(hipError_t e, const char* what);

class DeviceContext : public HostSpanSink {
 public:
  explicit DeviceContext(int device);
  ~DeviceContext();
  DeviceContext(const DeviceContext&) = delete;
  DeviceContext& operator=(const DeviceContext&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void* acquire(size_t bytes);
  void release(void* p, size_t bytes);
  void release_idle();
  uint64_t held() const { return held_; }
  uint64_t peak() const { return peak_; }
  // Scratch for per-launch partials, one growable buffer per slot (stream-ordered reuse).
  void* scratch(int slot, size_t bytes);
  // Pinned host memory for read-backs (one growable buffer).
  void* pinned(size_t bytes);
  // Coherent host memory that kernels write directly (results the host reads after a sync), one
  // growable buffer per slot (0..7): the host pointer, and the device pointer in `device`. Slots
  // 6 and 7 belong to the batched SyN loop.
  void* mapped(int slot, size_t bytes, void** device);
  // Pinned host space for a read-back whose value is wanted only at the next settle(): a bump
  // allocation valid until then (chunks are added, never moved, while copies may be in flight).
  void* stage(size_t bytes);
  // Waits for the stream and recycles the staging space.
  void settle();
  // A word of coherent host memory the device may store to (a hint channel: kernels report that
  // an iterative loop has stopped, so the host stops enqueueing its steps). Host and device views.
  volatile uint32_t* mailbox_host();
  uint32_t* mailbox_device();
  uint32_t next_tag() { return ++tag_; }
  // Completion words (in the mailbox): a kernel whose results go to coherent host memory stores
  // its tag into words [0, count) with a system-scope release once they are written, and
  // wait_done() spins until every word holds the tag: cheaper than a stream synchronisation.
  static constexpr int kDoneWords = 512;
  uint32_t* done_words_device() { return mailbox_device() + kDoneBase; }
  void wait_done(int count, uint32_t tag);
  void sync();
  // How long a wait on the device may take before it fails (device.rs:919-928
  // FASTWARP_GPU_WAIT_TIMEOUT_S: positive finite seconds, 300 by default), read at creation.
  double wait_timeout_s() const { return wait_timeout_s_; }
  // Launch-shape choices (the autotuner's table, loaded at creation; fw_set_tuning edits it).
  TuningTable& tuning() { return tuning_; }
  // Device copy of the Gaussian taps of (variance, axis length), cached.
  const float* taps(double variance, size_t length, int& radius);
  // The same taps on the host (f64), cached.
  const std::vector<double>& host_taps(double variance, size_t length);

  // Per-kernel timing (the reference's profiler.rs:95-170): while on, launches carry events;
  // take_profile() waits, adds the spans up per kernel family and clears them.
  void set_profiling(bool on);
  bool profiling() const { return profiling_; }
  // Limits the timing to every `every`-th launch of one kernel family (an empty name: every
  // launch of every family). The other launches are still counted, with their bytes, untimed, so
  // a timed run pays the events' cost on few launches.
  void set_profiling_focus(const std::string& family, int every);
  // Whether the next launch of `family` is to be timed (advances the focus family's counter).
  bool should_time(const char* family);
  hipEvent_t event();
  void drop(hipEvent_t e) { (void)hipEventDestroy(e); }
  void span(const char* name, double bytes, hipEvent_t a, hipEvent_t b);
  // An untimed launch of a family, with its algorithmic bytes.
  void count(const char* name, double bytes);
  void add_bytes(const char* name, double bytes);
  // Host work of the drivers while profiling, listed as "host:<name>" (sync waits included).
  void host_span(const char* name, double seconds) override;
  std::string take_profile();

 private:
  int device_;
  double wait_timeout_s_ = 300.0;
  TuningTable tuning_;
  hipStream_t stream_{};
  std::map<size_t, std::vector<void*>> free_;
  uint64_t held_ = 0, peak_ = 0;
  struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
  };
  Scratch scratch_[10];
  void* pinned_ = nullptr;
  struct StageChunk {
    char* p = nullptr;
    size_t capacity = 0, used = 0;
  };
  std::vector<StageChunk> stage_;
  struct Mapped {
    void* host = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
  };
  Mapped mapped_[8];
  static constexpr int kDoneBase = 64;  // words 0..63: the inversion's progress and stop words
  static constexpr size_t kMailboxBytes = sizeof(uint32_t) * (kDoneBase + kDoneWords);
  uint32_t* mailbox_ = nullptr;
  uint32_t* mailbox_dev_ = nullptr;
  uint32_t tag_ = 0;
  size_t pinned_bytes_ = 0;
  std::map<std::pair<double, size_t>, std::pair<void*, int>> taps_;
  std::map<std::pair<double, size_t>, std::vector<double>> host_taps_;
  bool profiling_ = false;
  std::string focus_;
  int focus_every_ = 1;
  uint64_t focus_seen_ = 0;
  struct Span {
    std::string name;
    double bytes;
    hipEvent_t a, b;
  };
  std::vector<Span> spans_;
  struct Total {
    uint64_t launches = 0, timed = 0;  // all launches, and those with events
    double ms = 0.0, bytes = 0.0;      // ms of the timed launches; bytes of all
  };
  std::map<std::string, Total> totals_;
};

// A timing scope: while the context profiles, the launches made inside it carry its events (or,
// for a family outside the profiling focus, are only counted).
class Timed {
 public:
  Timed(DeviceContext& ctx, const char* name, double bytes) : ctx_(ctx), name_(name), bytes_(bytes) {
    if (!ctx_.profiling()) return;
    active_ = true;
    if (!ctx_.should_time(name_)) return;
    ev_.start = ctx_.event();
    ev_.stop = ctx_.event();
    outer_ = current_launch_events();
    current_launch_events() = &ev_;
    timing_ = true;
  }
  ~Timed() {
    if (!active_) return;
    if (timing_) {
      current_launch_events() = outer_;
      if (ev_.launches > 0) {
        ctx_.span(name_, bytes_, ev_.start, ev_.stop);
        return;
      }
      ctx_.drop(ev_.start);
      ctx_.drop(ev_.stop);
    }
    ctx_.count(name_, bytes_);
  }
  Timed(const Timed&) = delete;
  Timed& operator=(const Timed&) = delete;

 private:
  DeviceContext& ctx_;
  const char* name_;
  double bytes_;
  bool active_ = false, timing_ = false;
  LaunchEvents ev_;
  LaunchEvents* outer_ = nullptr;
};

// A pooled device buffer of f32 values (move-only; returns to the pool on destruction).
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(DeviceContext* ctx, size_t floats);
  ~DevBuf() { reset(); }
  DevBuf(DevBuf&& o) noexcept : ctx_(o.ctx_), p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      ctx_ = o.ctx_;
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
    }
    return *this;
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  float* get() const { return p_; }
  size_t bytes() const { return bytes_; }
  explicit operator bool() const { return p_ != nullptr; }
  void reset();

 private:
  DeviceContext* ctx_ = nullptr;
  float* p_ = nullptr;
  size_t bytes_ = 0;
};

// The launch-shape table every context starts from: the file FASTWARP_TUNING names, else
// tuning/gfx950.json beside the library's directory (fastwarp_amd/tuning/ for lib/), else empty.
const TuningTable& default_tuning();

// A scalar image on the device (f32, x-fastest) with an upper bound of |value| (taken on the host
// at upload and carried through the ops that cannot grow it: warps, smoothing, resampling). The
// CC sums split every value at a power-of-two step of this bound.
struct HImage {
  Grid grid;
  DevBuf buf;
  float bound = 0.0f;
};

// A displacement field on the device: three f32 planes x, y, z.
struct HField {
  Grid grid;
  DevBuf buf;
};

using fw::InvertParams;
using fw::InvertStats;
using fw::MetricPair;

// The kernel arguments of a warp of an image on grid `src` through `field` (onto the field's
// grid), with an optional prior field and affine; the caller sets src, out (and the mask fields).
WarpArgs warp_args(const Grid& src, const HField& field, const HField* pre, const AffineTransform* affine);

class HipBackend {
 public:
  using ImageH = HImage;
  using FieldH = HField;

  explicit HipBackend(int device = 0);
  std::string name() const { return "hip-gfx950"; }
  DeviceContext& ctx() { return *ctx_; }

  void release_idle_buffers() { ctx_->release_idle(); }
  // Device bytes a new registration may use: FASTWARP_GPU_MEMORY_BUDGET_MB when set, else the
  // device's free memory less 1 GiB kept in reserve (capacity.rs:162-180).
  uint64_t memory_budget() const;
  HImage upload_image(const Image& v);
  HImage upload_image_f32(const Grid& g, const float* data);
  // register.rs:112-142: the image on the isotropic `mm` grid (same origin and direction,
  // floor((n - 1) * spacing / mm + 1e-9) + 1 voxels per axis), trilinear from the source indices
  // min(i * mm / spacing, n - 1); a copy when every axis already has spacing mm (within 1e-6).
  HImage resample_isotropic(const HImage& im, double mm);
  Image download_image(const HImage& im);
  HField zero_field(const Grid& g);
  HField upload_field(const VectorField& f);
  VectorField download_field(const HField& f);
  // Writes the field as 3n interleaved f64 values (x, y, z per voxel) into `aos`.
  void download_field_to(const HField& f, double* aos);
  Grid field_grid(const HField& f) const { return f.grid; }
  HField copy_field(const HField& f);

  HImage warp_through_field(const HImage& src, const HField& field, const HField* pre, const AffineTransform* affine);
  HImage warp_mask_through_field(const HImage& mask, const HField& field, const HField* pre,
                                 const AffineTransform* affine, HImage* product);

  // Both CC directions from one set of window sums; `valid` may be null; norms when want_norms.
  MetricPair<HField> cc_pair(const HImage& fixed, const HImage& moving, const HImage* valid, size_t radius,
                             bool want_norms);

  // Dense Mattes MI in both directions (masks may be null; bins 5..48).
  MetricPair<HField> mi_pair(const HImage& fixed, const HImage& moving, const HImage* fixed_mask, const HImage* moving_mask,
                             size_t bins, bool want_norms);

  // Level smoothing on the device (pyramid.rs:429-465, f32 recursive Gaussian, axes z, x, y);
  // sigma <= 0 gives a copy. The result's bound is its measured maximum magnitude.
  HImage smooth_image(const HImage& src, double sigma, SigmaUnits units);
  HImage smooth_image_mm(const HImage& src, const Vec3& sigma_mm);
  // Nearest-voxel and linear values at physical points (NaN outside), as the linear stages'
  // sample preparation reads them (mi.rs:161-207, linear.rs:251-279).
  void sample_values(const HImage& img, const HostVec<Vec3>& points, HostVec<double>& nearest,
                     HostVec<double>& linear);
  // Finite minimum and maximum; false when the image has no finite value.
  bool finite_range(const HImage& img, double& lo, double& hi);
  // Both sides of a SyN step in single launches (the SyN driver's pair operations): two warps onto
  // one grid, two compositions, two update smoothings. Each result equals the single operation's.
  std::pair<HImage, HImage> warp_pair(const HImage& src0, const HField& field0, const AffineTransform* affine0,
                                      const HImage& src1, const HField& field1, const AffineTransform* affine1);
  std::pair<HField, HField> compose_pair(const HField& u0, const HField& w0, const HField& u1, const HField& w1);
  std::pair<HField, HField> smooth_field_pair(HField a, HField b, double variance);
  // The winsorised, windowed copy of an image already in HBM (preprocess.rs:159-232), raising the
  // registration's error for non-finite voxels; and its intensity-weighted centre (init.rs:47-86).
  HImage preprocess_image(const HImage& raw, float lower, float upper, double out_min = 0.0, double out_max = 1.0);
  Vec3 centre_of_gravity(const HImage& img);

  HField smooth_field(HField field, double variance);
  std::pair<double, double> max_scaled_norm_pair(const HField& a, const HField* b);
  double l2_norm(const HField& f);
  void scale(HField& f, double s);
  void axpy(HField& acc, double a, const HField& x);
  HField compose(const HField& displacement, const HField& warping);
  std::pair<HField, InvertStats> invert(const HField& field, HField initial, const InvertParams& p);

  // Deferred forms for the SyN loop: results read back at the next settle().
  MetricPair<HField> cc_pair_deferred(const HImage& fixed, const HImage& moving, size_t radius);
  // Scales both updates so each one's largest scaled voxel norm is `step` (step / max, or step when
  // the maximum is not positive), on the device; the two maxima come back at settle().
  Pending<std::pair<double, double>> scale_to_step(HField& a, HField& b, double step);
  // The same factors without scaling the fields (f32 factors on the device, the maxima at
  // settle()), for compose_pair_scaled to apply while it interpolates the updates.
  std::pair<Pending<std::pair<double, double>>, DevBuf> step_factors(const HField& a, const HField& b, double step);
  // compose_pair with the updates u0, u1 multiplied by factors[0], factors[1] (device).
  std::pair<HField, HField> compose_pair_scaled(const HField& u0, const HField& w0, const HField& u1, const HField& w1,
                                                const DevBuf& factors);
  std::pair<HField, Pending<InvertStats>> invert_deferred(const HField& field, HField initial, const InvertParams& p);
  // invert_deferred of two fields on one grid in the same launches (each item as alone, so the
  // same bits): {first estimate, second estimate, their statistics}.
  struct InvertedPair {
    HField first, second;
    Pending<InvertStats> first_stats, second_stats;
  };
  InvertedPair invert_pair_deferred(const HField& f0, HField initial0, const HField& f1, HField initial1,
                                    const InvertParams& p);
  void settle() { ctx_->settle(); }
  HField resample_field(const HField& field, const Grid& grid);

  // CC tile height in z of the fused kernel.
  int cc_tile_z = kCcTileZ;
  // Slots of the coherent-host rings of update maxima and CC partials (results awaiting a settle).
  static constexpr size_t kMaximaSlots = 64;
  static constexpr size_t kCcSlots = 4;
  static constexpr size_t kCcSlotBlocks = size_t(1) << 16;

 private:
  std::shared_ptr<DeviceContext> ctx_;
};

// The linear-stage metric of one level on the device (the reference's GpuMiLevel,
// crates/fastwarp-gpu/src/wgpu_backend/mi.rs:144-524): sample records resident in HBM, one
// histogram pass per value, plus one derivative pass with FP64 block reductions.
class MiLevel {
 public:
  // `moving` must outlive the level. points: physical sample points; fixed_bin: their bins.
  MiLevel(HipBackend& backend, const HImage& moving, const HostVec<Vec3>& points, const HostVec<uint32_t>& fixed_bin,
          const MiBinning& binning);
  double value(const LinearMovingTransform& t);
  // The values of several transforms, batched into as few histogram passes as the LDS allows.
  std::vector<double> values(const std::vector<LinearMovingTransform>& ts);
  // Whether batching pays on this level: below about half a million samples a pass is bound by
  // its latency, and an extra transform in it costs a few microseconds against a round trip of
  // some twenty; on larger levels every transform costs as much as a pass of its own.
  bool batch_pays() const { return count_ <= kBatchPaysSamples && mi_batch_capacity(int(binning_.bins)) > 1; }
  static constexpr size_t kBatchPaysSamples = 500000;
  std::pair<double, std::vector<double>> value_and_derivative(const LinearMovingTransform& t);
  // The histogram of one transform as the finish reads it: the joint cells (bins x bins, in
  // samples: the B-spline weights summed), the fixed marginal counts and the valid count.
  void histogram_cells(const LinearMovingTransform& t, std::vector<double>& joint, std::vector<double>& fixed_marginal,
                       uint64_t& valid);
  size_t count() const { return count_; }

 private:
  void batch_values(const LinearMovingTransform* ts, size_t k, std::vector<double>& out);
  int pass_args(const LinearMovingTransform& t, MiSampledArgs& args, size_t histograms);
  void launch_finish(const unsigned long long* slots, int blocks, size_t k, float* ratio, MiFinishOut* out, bool signal,
                     uint32_t tag);
  volatile MiFinishOut* finish_slots(size_t k, MiFinishOut** dev);
  void evaluate(const MiSampledArgs& a, const MiBatch& batch, size_t k, int blocks, float* ratio, MiFinishOut* out, bool signal,
                uint32_t tag);
  DeviceContext* ctx_;
  const HImage* moving_;
  DevBuf samples_;
  size_t count_ = 0;
  Vec3 anchor_{0, 0, 0};
  MiBinning binning_;
};

// A LinearEngine (linear.rs:288-320) whose levels share the device-resident smoothed moving image.
class HipLinearEngine {
 public:
  explicit HipLinearEngine(HipBackend& b) : backend_(b) {}
  struct Level {
    std::shared_ptr<const HImage> moving;
    std::unique_ptr<MiLevel> metric;
    double value(const LinearMovingTransform& t) { return metric->value(t); }
    std::vector<double> values(const std::vector<LinearMovingTransform>& ts) { return metric->values(ts); }
    bool batch_pays() const { return metric->batch_pays(); }
    std::pair<double, std::vector<double>> value_and_derivative(const LinearMovingTransform& t) {
      return metric->value_and_derivative(t);
    }
  };
  Level prepare_level(std::shared_ptr<const HImage> moving, const HostVec<Vec3>& points,
                      const HostVec<uint32_t>& fixed_bin, const MiBinning& binning) {
    Level l;
    l.moving = std::move(moving);
    l.metric = std::make_unique<MiLevel>(backend_, *l.moving, points, fixed_bin, binning);
    return l;
  }
  std::string name() const { return "hip-gfx950"; }

 private:
  HipBackend& backend_;
};

}  // namespace fw::hip
