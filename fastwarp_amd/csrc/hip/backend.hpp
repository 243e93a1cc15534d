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

#include "../host/grid.hpp"
#include "../host/transform.hpp"
#include "kernels.hpp"

namespace fw::hip {

void check(hipError_t e, const char* what);

class DeviceContext {
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
  void sync();
  // Device copy of the Gaussian taps of (variance, axis length), cached.
  const float* taps(double variance, size_t length, int& radius);

 private:
  int device_;
  hipStream_t stream_{};
  std::map<size_t, std::vector<void*>> free_;
  uint64_t held_ = 0, peak_ = 0;
  struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
  };
  Scratch scratch_[8];
  void* pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
  std::map<std::pair<double, size_t>, std::pair<void*, int>> taps_;
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

// A scalar image on the device (f32, x-fastest) with an upper bound of |value| (for the
// fixed-point scale of the CC sums; warps of an image keep its bound).
struct HImage {
  Grid grid;
  DevBuf buf;
  float absmax = 0.0f;
};

// A displacement field on the device: three f32 planes x, y, z.
struct HField {
  Grid grid;
  DevBuf buf;
};

struct InvertParams {
  uint32_t max_iterations = 20;
  double mean_error_tolerance = 0.001;
  double max_error_tolerance = 0.1;
  bool enforce_boundary_condition = true;
};

struct InvertStats {
  uint32_t iterations = 0;
  double mean_error_norm = 0.0;
  double max_error_norm = 0.0;
};

template <typename F>
struct MetricPair {
  double values[2] = {0, 0};
  F fields[2];
  double norms[2] = {0, 0};
};

class HipBackend {
 public:
  using ImageH = HImage;
  using FieldH = HField;

  explicit HipBackend(int device = 0);
  std::string name() const { return "hip-gfx950"; }
  DeviceContext& ctx() { return *ctx_; }

  void release_idle_buffers() { ctx_->release_idle(); }
  HImage upload_image(const Image& v);
  HImage upload_image_f32(const Grid& g, const float* data);
  Image download_image(const HImage& im);
  HField zero_field(const Grid& g);
  HField upload_field(const VectorField& f);
  VectorField download_field(const HField& f);
  Grid field_grid(const HField& f) const { return f.grid; }
  HField copy_field(const HField& f);

  HImage warp_through_field(const HImage& src, const HField& field, const HField* pre, const AffineTransform* affine);
  HImage warp_mask_through_field(const HImage& mask, const HField& field, const HField* pre,
                                 const AffineTransform* affine, HImage* product);

  // Both CC directions from one set of window sums; `valid` may be null; norms when want_norms.
  MetricPair<HField> cc_pair(const HImage& fixed, const HImage& moving, const HImage* valid, size_t radius,
                             bool want_norms);

  HField smooth_field(HField field, double variance);
  std::pair<double, double> max_scaled_norm_pair(const HField& a, const HField* b);
  double l2_norm(const HField& f);
  void scale(HField& f, double s);
  void axpy(HField& acc, double a, const HField& x);
  HField compose(const HField& displacement, const HField& warping);
  std::pair<HField, InvertStats> invert(const HField& field, HField initial, const InvertParams& p);
  HField resample_field(const HField& field, const Grid& grid);

  // CC tile height in z of the fused kernel.
  int cc_tile_z = 16;

 private:
  std::shared_ptr<DeviceContext> ctx_;
};

}  // namespace fw::hip
