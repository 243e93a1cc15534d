// HipBackend implementation: buffer pool, host<->device conversions and the mapping of every
// trait op (crates/fastwarp-reg/src/backend.rs) onto the kernels of kernels.hip. Geometry is
// composed in f64 on the host into continuous-index maps and narrowed to f32 once per launch.
#include "backend.hpp"

#include <dlfcn.h>

#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <thread>

#include "../host/gauss_taps.hpp"
#include "../host/preprocess.hpp"
#include "geometry.hpp"

namespace fw::hip {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error("hip", std::string(what) + ": " + hipGetErrorString(e));
}

const TuningTable& default_tuning() {
  static const TuningTable table = [] {
    TuningTable t;
    if (const char* env = std::getenv("FASTWARP_TUNING")) {
      t.load(env);
      return t;
    }
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&default_tuning), &info) && info.dli_fname) {
      std::string dir(info.dli_fname);
      const size_t slash = dir.rfind('/');
      dir = slash == std::string::npos ? "." : dir.substr(0, slash);
      t.load(dir + "/../tuning/gfx950.json");
    }
    return t;
  }();
  return table;
}

DeviceContext::DeviceContext(int device) : device_(device), tuning_(default_tuning()) {
  wait_timeout_s_ = wait_timeout_from(std::getenv("FASTWARP_GPU_WAIT_TIMEOUT_S"));
  check(hipSetDevice(device), "hipSetDevice");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}

DeviceContext::~DeviceContext() {
  if (host_span_sink() == this) host_span_sink() = nullptr;
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  release_idle();
  for (auto& s : scratch_)
    if (s.p) hipFree(s.p);
  if (pinned_) hipHostFree(pinned_);
  if (mailbox_) hipHostFree(mailbox_);
  for (auto& m : mapped_)
    if (m.host) hipHostFree(m.host);
  for (auto& c : stage_) hipHostFree(c.p);
  for (auto& kv : taps_) hipFree(kv.second.first);
  for (auto& s : spans_) {
    (void)hipEventDestroy(s.a);
    (void)hipEventDestroy(s.b);
  }
  (void)hipStreamDestroy(stream_);
}

static size_t round_bytes(size_t b) { return (b + 255) & ~size_t(255); }

void* DeviceContext::acquire(size_t bytes) {
  bytes = round_bytes(std::max<size_t>(bytes, 256));
  auto it = free_.find(bytes);
  if (it != free_.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    return p;
  }
  (void)hipSetDevice(device_);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipStreamSynchronize(stream_);
    release_idle();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error("hip", "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
  }
  held_ += bytes;
  peak_ = std::max(peak_, held_);
  return p;
}

void DeviceContext::release(void* p, size_t bytes) {
  if (!p) return;
  free_[round_bytes(std::max<size_t>(bytes, 256))].push_back(p);
}

void DeviceContext::release_idle() {
  (void)hipStreamSynchronize(stream_);
  for (auto& kv : free_)
    for (void* p : kv.second) {
      hipFree(p);
      held_ -= kv.first;
    }
  free_.clear();
}

void* DeviceContext::scratch(int slot, size_t bytes) {
  Scratch& s = scratch_[slot];
  if (s.bytes < bytes) {
    if (s.p) {
      hipStreamSynchronize(stream_);
      hipFree(s.p);
    }
    s.bytes = round_bytes(bytes * 2);
    check(hipMalloc(&s.p, s.bytes), "hipMalloc scratch");
  }
  return s.p;
}

void* DeviceContext::pinned(size_t bytes) {
  if (pinned_bytes_ < bytes) {
    if (pinned_) hipHostFree(pinned_);
    pinned_bytes_ = round_bytes(bytes * 2);
    check(hipHostMalloc(&pinned_, pinned_bytes_, hipHostMallocDefault), "hipHostMalloc");
  }
  return pinned_;
}

void* DeviceContext::mapped(int slot, size_t bytes, void** device) {
  Mapped& m = mapped_[slot];
  if (m.bytes < bytes) {
    if (m.host) {
      sync();
      hipHostFree(m.host);
    }
    m.bytes = round_bytes(bytes * 2);
    check(hipHostMalloc(&m.host, m.bytes, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc mapped");
    check(hipHostGetDevicePointer(&m.dev, m.host, 0), "hipHostGetDevicePointer");
  }
  *device = m.dev;
  return m.host;
}

void* DeviceContext::stage(size_t bytes) {
  bytes = round_bytes(bytes);
  for (auto& c : stage_)
    if (c.capacity - c.used >= bytes) {
      void* p = c.p + c.used;
      c.used += bytes;
      return p;
    }
  StageChunk c;
  c.capacity = std::max<size_t>(bytes, size_t(1) << 20);
  void* p = nullptr;
  check(hipHostMalloc(&p, c.capacity, hipHostMallocDefault), "hipHostMalloc stage");
  c.p = static_cast<char*>(p);
  c.used = bytes;
  stage_.push_back(c);
  return c.p;
}

void DeviceContext::settle() {
  sync();
  for (auto& c : stage_) c.used = 0;
}

volatile uint32_t* DeviceContext::mailbox_host() {
  if (!mailbox_) {
    void* p = nullptr;
    check(hipHostMalloc(&p, kMailboxBytes, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc mailbox");
    mailbox_ = static_cast<uint32_t*>(p);
    std::memset(mailbox_, 0, kMailboxBytes);
    void* d = nullptr;
    check(hipHostGetDevicePointer(&d, p, 0), "hipHostGetDevicePointer");
    mailbox_dev_ = static_cast<uint32_t*>(d);
  }
  return mailbox_;
}

uint32_t* DeviceContext::mailbox_device() {
  (void)mailbox_host();
  return mailbox_dev_;
}

// Waits for the stream by polling it, so that a device that stops making progress fails the call
// after the wait timeout instead of blocking it for good (a device error is reported at once).
static void wait_stream(hipStream_t stream, std::chrono::steady_clock::time_point t0, double limit_s) {
  for (unsigned polls = 0;; ++polls) {
    const hipError_t e = hipStreamQuery(stream);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) check(e, "hipStreamQuery");
    if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > limit_s)
      invalid("hip", "the device did not finish within " + std::to_string(limit_s) + " s (FASTWARP_GPU_WAIT_TIMEOUT_S)");
    if (polls > 1000) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void DeviceContext::wait_done(int count, uint32_t tag) {
  HostTimer t("sync_wait");
  if (count < 1 || count > kDoneWords) invalid("device", "completion word count outside 1.." + std::to_string(kDoneWords));
  volatile uint32_t* w = mailbox_host() + kDoneBase;
  const auto t0 = std::chrono::steady_clock::now();
  for (int b = 0; b < count; ++b) {
    for (unsigned spins = 1; w[b] != tag; ++spins) {
      if ((spins & 4095) != 0) continue;
      // now and then: a device error is reported, a stream that has drained without the word is
      // an error, and the wait fails after the wait timeout
      const hipError_t e = hipStreamQuery(stream_);
      if (e != hipSuccess && e != hipErrorNotReady) check(e, "hipStreamQuery");
      if (e == hipSuccess && w[b] != tag) invalid("device", "a kernel completed without writing its completion word");
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > wait_timeout_s_)
        invalid("hip", "the device did not finish within " + std::to_string(wait_timeout_s_) +
                           " s (FASTWARP_GPU_WAIT_TIMEOUT_S)");
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

void DeviceContext::sync() {
  HostTimer t("sync_wait");
  wait_stream(stream_, std::chrono::steady_clock::now(), wait_timeout_s_);
}

const std::vector<double>& DeviceContext::host_taps(double variance, size_t length) {
  auto key = std::make_pair(variance, length);
  auto it = host_taps_.find(key);
  if (it == host_taps_.end()) it = host_taps_.emplace(key, gaussian_operator_taps(variance, 0.001, length)).first;
  return it->second;
}

const float* DeviceContext::taps(double variance, size_t length, int& radius) {
  auto key = std::make_pair(variance, length);
  auto it = taps_.find(key);
  if (it == taps_.end()) {
    const std::vector<double> t = gaussian_operator_taps(variance, 0.001, length);
    std::vector<float> f(t.begin(), t.end());
    void* p = nullptr;
    check(hipMalloc(&p, f.size() * sizeof(float)), "hipMalloc taps");
    check(hipMemcpy(p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy taps");
    it = taps_.emplace(key, std::make_pair(p, int(t.size() / 2))).first;
  }
  radius = it->second.second;
  return static_cast<const float*>(it->second.first);
}

void DeviceContext::set_profiling(bool on) {
  if (!on && profiling_) (void)take_profile();
  profiling_ = on;
  HostSpanSink*& sink = host_span_sink();
  if (on) sink = this;
  else if (sink == this) sink = nullptr;
}

void DeviceContext::host_span(const char* name, double seconds) {
  Total& t = totals_[std::string("host:") + name];
  t.launches += 1;
  t.timed += 1;
  t.ms += seconds * 1e3;
}

// Every span gets new events: an event that has been carried by a launch before is not handed
// to another (its earlier binding is not relied on to be replaced).
hipEvent_t DeviceContext::event() {
  hipEvent_t e = nullptr;
  check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

void DeviceContext::set_profiling_focus(const std::string& family, int every) {
  if (every < 1) invalid("profiling", "the timing interval must be at least 1");
  focus_ = family;
  focus_every_ = every;
  focus_seen_ = 0;
}

bool DeviceContext::should_time(const char* family) {
  if (focus_.empty()) return true;
  if (focus_ != family) return false;
  return focus_seen_++ % uint64_t(focus_every_) == 0;
}

void DeviceContext::span(const char* name, double bytes, hipEvent_t a, hipEvent_t b) { spans_.push_back({name, bytes, a, b}); }

void DeviceContext::count(const char* name, double bytes) {
  Total& t = totals_[name];
  t.launches += 1;
  t.bytes += bytes;
}

void DeviceContext::add_bytes(const char* name, double bytes) { totals_[name].bytes += bytes; }

std::string DeviceContext::take_profile() {
  sync();
  for (Span& s : spans_) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, s.a, s.b), "hipEventElapsedTime");
    Total& t = totals_[s.name];
    t.launches += 1;
    t.timed += 1;
    t.ms += double(ms);
    t.bytes += s.bytes;
    (void)hipEventDestroy(s.a);
    (void)hipEventDestroy(s.b);
  }
  spans_.clear();
  std::string out;
  char line[256];
  for (auto& kv : totals_) {
    std::snprintf(line, sizeof line, "%s %llu %.6f %.0f %llu\n", kv.first.c_str(), (unsigned long long)kv.second.launches,
                  kv.second.ms, kv.second.bytes, (unsigned long long)kv.second.timed);
    out += line;
  }
  totals_.clear();
  return out;
}

DevBuf::DevBuf(DeviceContext* ctx, size_t floats) : ctx_(ctx), bytes_(floats * sizeof(float)) {
  p_ = static_cast<float*>(ctx->acquire(bytes_));
}

void DevBuf::reset() {
  if (p_ && ctx_) ctx_->release(p_, bytes_);
  p_ = nullptr;
}

// ---------------------------------------------------------------------------- geometry ---


HipBackend::HipBackend(int device) : ctx_(std::make_shared<DeviceContext>(device)) {}

// --------------------------------------------------------------------- transfers (a1, a2) ---

uint64_t HipBackend::memory_budget() const {
  if (const char* env = std::getenv("FASTWARP_GPU_MEMORY_BUDGET_MB")) {
    char* end = nullptr;
    const unsigned long long mb = std::strtoull(env, &end, 10);
    if (end != env) return uint64_t(mb) << 20;
  }
  size_t free_bytes = 0, total = 0;
  check(hipMemGetInfo(&free_bytes, &total), "hipMemGetInfo");
  const uint64_t reserve = uint64_t(1) << 30;
  return free_bytes > reserve ? uint64_t(free_bytes) - reserve : 0;
}

HImage HipBackend::upload_image(const Image& v) {
  HostTimer host_timer("upload_image");
  const size_t n = v.grid.num_voxels();
  auto* stage = static_cast<float*>(ctx_->pinned(n * sizeof(float)));
  constexpr size_t kGrain = size_t(1) << 16;
  std::vector<float> bound((n + kGrain - 1) / kGrain, 0.0f);
  parallel_chunks(n, kGrain, [&](size_t c, size_t b, size_t e) {
    float m = 0.0f;
    for (size_t i = b; i < e; ++i) {
      const float x = float(v.data[i]);
      stage[i] = x;
      if (std::isfinite(x)) m = std::max(m, std::fabs(x));
    }
    bound[c] = m;
  });
  float m = 0.0f;
  for (float x : bound) m = std::max(m, x);
  HImage im{v.grid, DevBuf(ctx_.get(), n), m};
  check(hipMemcpyAsync(im.buf.get(), stage, n * sizeof(float), hipMemcpyHostToDevice, ctx_->stream()), "upload image");
  ctx_->sync();
  return im;
}

HImage HipBackend::resample_isotropic(const HImage& im, double mm) {
  if (!(std::isfinite(mm) && mm > 0.0)) invalid("register", "isotropic spacing must be positive");
  const Grid& g = im.grid;
  const size_t n = g.num_voxels();
  if (std::fabs(g.spacing[0] - mm) < 1e-6 && std::fabs(g.spacing[1] - mm) < 1e-6 && std::fabs(g.spacing[2] - mm) < 1e-6) {
    HImage copy{g, DevBuf(ctx_.get(), n), im.bound};
    check(hipMemcpyAsync(copy.buf.get(), im.buf.get(), n * sizeof(float), hipMemcpyDeviceToDevice, ctx_->stream()), "copy image");
    return copy;
  }
  // the source index of every output row, column and plane, split in f64 as the reference does
  std::array<size_t, 3> out{1, 1, 1};
  std::vector<int> lo;
  std::vector<float> fr;
  for (int a = 0; a < 3; ++a) {
    const double last = double(g.size[a] - 1);
    const double ratio = mm / g.spacing[a];
    out[a] = std::max<size_t>(size_t(std::floor(last * g.spacing[a] / mm + 1e-9)) + 1, 1);
    for (size_t i = 0; i < out[a]; ++i) {
      const double c = std::min(double(i) * ratio, last);
      const double f = std::floor(c);
      lo.push_back(int(f));
      fr.push_back(float(c - f));
    }
  }
  const Grid og(out, {mm, mm, mm}, g.origin, g.direction);
  DevBuf tables(ctx_.get(), lo.size() + fr.size());
  check(hipMemcpyAsync(tables.get(), lo.data(), lo.size() * sizeof(int), hipMemcpyHostToDevice, ctx_->stream()), "upload table");
  check(hipMemcpyAsync(tables.get() + lo.size(), fr.data(), fr.size() * sizeof(float), hipMemcpyHostToDevice, ctx_->stream()),
        "upload table");
  HImage r{og, DevBuf(ctx_.get(), og.num_voxels()), im.bound};
  {
    // the source read once (at most), the result written
    Timed t(*ctx_, "k_resample_iso", 4.0 * double(og.num_voxels()) + 4.0 * std::min(double(n), double(og.num_voxels())));
    launch_resample_iso(im.buf.get(), dims_of(g), r.buf.get(), dims_of(og), reinterpret_cast<const int*>(tables.get()),
                        tables.get() + lo.size(), ctx_->stream());
  }
  // the tables are host vectors: the copies must finish before they go
  ctx_->sync();
  return r;
}

HImage HipBackend::upload_image_f32(const Grid& g, const float* data) {
  const size_t n = g.num_voxels();
  constexpr size_t kGrain = size_t(1) << 16;
  std::vector<float> bound((n + kGrain - 1) / kGrain, 0.0f);
  parallel_chunks(n, kGrain, [&](size_t c, size_t b, size_t e) {
    float m = 0.0f;
    for (size_t i = b; i < e; ++i)
      if (std::isfinite(data[i])) m = std::max(m, std::fabs(data[i]));
    bound[c] = m;
  });
  float m = 0.0f;
  for (float x : bound) m = std::max(m, x);
  HImage im{g, DevBuf(ctx_.get(), n), m};
  check(hipMemcpyAsync(im.buf.get(), data, n * sizeof(float), hipMemcpyHostToDevice, ctx_->stream()), "upload image");
  ctx_->sync();
  return im;
}

Image HipBackend::download_image(const HImage& im) {
  const size_t n = im.grid.num_voxels();
  std::vector<float> f(n);
  check(hipMemcpyAsync(f.data(), im.buf.get(), n * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "download image");
  ctx_->sync();
  return Image(im.grid, std::vector<double>(f.begin(), f.end()));
}

HField HipBackend::zero_field(const Grid& g) {
  HField f{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  check(hipMemsetAsync(f.buf.get(), 0, f.buf.bytes(), ctx_->stream()), "zero field");
  return f;
}

HField HipBackend::upload_field(const VectorField& v) {
  HostTimer host_timer("upload_field");
  const size_t n = v.grid.num_voxels();
  std::vector<float> soa(3 * n);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i)
      for (int c = 0; c < 3; ++c) soa[c * n + i] = float(v.data[i][c]);
  });
  HField f{v.grid, DevBuf(ctx_.get(), 3 * n)};
  check(hipMemcpyAsync(f.buf.get(), soa.data(), soa.size() * sizeof(float), hipMemcpyHostToDevice, ctx_->stream()),
        "upload field");
  ctx_->sync();
  return f;
}

void HipBackend::download_field_to(const HField& f, double* aos) {
  HostTimer host_timer("download_field");
  const size_t n = f.grid.num_voxels();
  auto* stage = static_cast<float*>(ctx_->pinned(3 * n * sizeof(float)));
  check(hipMemcpyAsync(stage, f.buf.get(), 3 * n * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "download field");
  ctx_->sync();
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      aos[3 * i] = double(stage[i]);
      aos[3 * i + 1] = double(stage[n + i]);
      aos[3 * i + 2] = double(stage[2 * n + i]);
    }
  });
}

VectorField HipBackend::download_field(const HField& f) {
  HostTimer host_timer("download_field");
  const size_t n = f.grid.num_voxels();
  std::vector<float> soa(3 * n);
  check(hipMemcpyAsync(soa.data(), f.buf.get(), soa.size() * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()),
        "download field");
  ctx_->sync();
  std::vector<Vec3> aos(n);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) aos[i] = {double(soa[i]), double(soa[n + i]), double(soa[2 * n + i])};
  });
  return VectorField(f.grid, std::move(aos));
}

HField HipBackend::copy_field(const HField& f) {
  HField out{f.grid, DevBuf(ctx_.get(), 3 * f.grid.num_voxels())};
  check(hipMemcpyAsync(out.buf.get(), f.buf.get(), f.buf.bytes(), hipMemcpyDeviceToDevice, ctx_->stream()), "copy field");
  return out;
}

// ------------------------------------------------------------------------ warps (a3, a4) ---

WarpArgs warp_args(const Grid& src, const HField& field, const HField* pre, const AffineTransform* affine) {
  WarpArgs a{};
  const Grid& og = field.grid;
  const Mat3 M = affine ? affine->matrix : kIdentity3;
  const Vec3 off = affine ? affine->offset() : Vec3{0, 0, 0};
  const Mat3 L = mat3_mul(src.index_from_offset, M);
  // c = IFO_s (M (x + u + d) + off - o_s), x = o_f + D_f S_f idx
  a.to_src = {};
  put(a.to_src.m, mat3_mul(L, scaled_direction(og)));
  put(a.to_src.c0, mat3_mul_vec(src.index_from_offset, vsub(vadd(mat3_mul_vec(M, og.origin), off), src.origin)));
  put(a.to_src.l, L);
  put(a.pre_l, L);
  a.sd = dims_of(src);
  a.od = dims_of(og);
  a.field = field.buf.get();
  if (pre) {
    const Grid& pg = pre->grid;
    put(a.to_pre.m, mat3_mul(pg.index_from_offset, scaled_direction(og)));
    put(a.to_pre.c0, mat3_mul_vec(pg.index_from_offset, vsub(og.origin, pg.origin)));
    put(a.to_pre.l, pg.index_from_offset);
    a.pre = pre->buf.get();
    a.pd = dims_of(pg);
  }
  return a;
}

HImage HipBackend::warp_through_field(const HImage& src, const HField& field, const HField* pre, const AffineTransform* affine) {
  WarpArgs a = warp_args(src.grid, field, pre, affine);
  HImage out{field.grid, DevBuf(ctx_.get(), field.grid.num_voxels()), src.bound};
  a.src = src.buf.get();
  a.out = out.buf.get();
  const double no = double(field.grid.num_voxels());
  Timed t(*ctx_, "k_warp", (16.0 + (pre ? 12.0 : 0.0)) * no + 4.0 * std::min(no, double(src.grid.num_voxels())));
  launch_warp(a, ctx_->stream());
  return out;
}

std::pair<HImage, HImage> HipBackend::warp_pair(const HImage& src0, const HField& field0, const AffineTransform* affine0,
                                                const HImage& src1, const HField& field1, const AffineTransform* affine1) {
  if (field0.grid != field1.grid)
    return {warp_through_field(src0, field0, nullptr, affine0), warp_through_field(src1, field1, nullptr, affine1)};
  WarpArgs a0 = warp_args(src0.grid, field0, nullptr, affine0), a1 = warp_args(src1.grid, field1, nullptr, affine1);
  const Grid& g = field0.grid;
  HImage out0{g, DevBuf(ctx_.get(), g.num_voxels()), src0.bound}, out1{g, DevBuf(ctx_.get(), g.num_voxels()), src1.bound};
  a0.src = src0.buf.get();
  a0.out = out0.buf.get();
  a1.src = src1.buf.get();
  a1.out = out1.buf.get();
  const double no = double(g.num_voxels());
  Timed t(*ctx_, "k_warp", 32.0 * no + 4.0 * (std::min(no, double(src0.grid.num_voxels())) + std::min(no, double(src1.grid.num_voxels()))));
  launch_warp2(a0, a1, ctx_->stream());
  return {std::move(out0), std::move(out1)};
}

HImage HipBackend::warp_mask_through_field(const HImage& mask, const HField& field, const HField* pre,
                                           const AffineTransform* affine, HImage* product) {
  if (product && product->grid != field.grid) invalid("warp mask", "the product image is not on the grid of the field");
  WarpArgs a = warp_args(mask.grid, field, pre, affine);
  a.src = mask.buf.get();
  a.nearest = 1;
  if (product) {
    HImage out = std::move(*product);
    out.bound *= mask.bound;
    a.product = out.buf.get();
    a.out = out.buf.get();
    Timed t(*ctx_, "k_warp", 24.0 * double(field.grid.num_voxels()));
    launch_warp(a, ctx_->stream());
    return out;
  }
  HImage out{field.grid, DevBuf(ctx_.get(), field.grid.num_voxels()), mask.bound};
  a.out = out.buf.get();
  Timed t(*ctx_, "k_warp", 20.0 * double(field.grid.num_voxels()));
  launch_warp(a, ctx_->stream());
  return out;
}

// ----------------------------------------------------------------------------- CC (a5, a6) ---


MetricPair<HField> HipBackend::cc_pair(const HImage& fixed, const HImage& moving, const HImage* valid, size_t radius,
                                       bool want_norms) {
  if (fixed.grid != moving.grid) invalid("hip backend", "cc: the images are not on one grid");
  if (valid && valid->grid != fixed.grid) invalid("hip backend", "cc mask: the images are not on one grid");
  if (radius == 0 || radius > 8) invalid("hip backend", "cc radius " + std::to_string(radius) + " outside 1..=8");
  const Grid& g = fixed.grid;
  const size_t n = g.num_voxels();
  MetricPair<HField> out;
  out.fields[0] = HField{g, DevBuf(ctx_.get(), 3 * n)};
  out.fields[1] = HField{g, DevBuf(ctx_.get(), 3 * n)};
  CcArgs a{};
  a.fixed = fixed.buf.get();
  a.moving = moving.buf.get();
  a.valid = valid ? valid->buf.get() : nullptr;
  a.d = dims_of(g);
  a.radius = int(radius);
  a.tile_z = cc_tile_for(a.d, int(ctx_->tuning().get("cc.tile_z", n, cc_tile_z)),
                         int(ctx_->tuning().get("cc.min_blocks", n, 512)));
  a.scale_f = cc_split_scale(fixed.bound);
  a.scale_m = cc_split_scale(moving.bound);
  a.step_f = 1.0f / a.scale_f;
  a.step_m = 1.0f / a.scale_m;
  a.probes = probes_of(g);
  a.d_fixed = out.fields[0].buf.get();
  a.d_moving = out.fields[1].buf.get();
  const dim3 gr = cc_grid(a.d, a.tile_z);
  const size_t blocks = size_t(gr.x) * gr.y * gr.z;
  a.partial_cc = static_cast<float*>(ctx_->scratch(0, blocks * sizeof(float)));
  a.partial_count = static_cast<int*>(ctx_->scratch(1, blocks * sizeof(int)));
  {
    // F and M read, both derivative fields written (the mask too when there is one)
    Timed t(*ctx_, "k_cc2", (valid ? 36.0 : 32.0) * double(n));
    launch_cc_pair(a, ctx_->stream());
  }
  const size_t real = blocks;
  int nb = 0;
  float* pm = nullptr;
  float* ps = nullptr;
  if (want_norms) {
    nb = norm_blocks(long(n));
    pm = static_cast<float*>(ctx_->scratch(2, 2 * nb * sizeof(float)));
    ps = static_cast<float*>(ctx_->scratch(3, 2 * nb * sizeof(float)));
    Timed t(*ctx_, "k_field_norms", 24.0 * double(n));
    launch_field_norms(out.fields[0].buf.get(), out.fields[1].buf.get(), long(n), inv_spacing(g), pm, ps, nb, ctx_->stream());
  }
  const size_t bytes = real * (sizeof(float) + sizeof(int)) + (want_norms ? 2 * nb * sizeof(float) : 0);
  char* host = static_cast<char*>(ctx_->pinned(bytes));
  check(hipMemcpyAsync(host, a.partial_cc, real * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "cc partials");
  check(hipMemcpyAsync(host + real * sizeof(float), a.partial_count, real * sizeof(int), hipMemcpyDeviceToHost, ctx_->stream()),
        "cc counts");
  if (want_norms)
    check(hipMemcpyAsync(host + real * (sizeof(float) + sizeof(int)), ps, 2 * nb * sizeof(float), hipMemcpyDeviceToHost,
                         ctx_->stream()),
          "cc norms");
  ctx_->sync();
  const float* pcc = reinterpret_cast<const float*>(host);
  const int* pcount = reinterpret_cast<const int*>(host + real * sizeof(float));
  double total = 0.0;
  long long count = 0;
  for (size_t b = 0; b < real; ++b) {
    total += double(pcc[b]);
    count += pcount[b];
  }
  if (count == 0)
    invalid("cc metric", valid ? "no voxel of the registration level lies inside both the fixed and the moving mask"
                               : "no virtual voxel maps inside both image buffers");
  out.values[0] = out.values[1] = -total / double(count);
  if (want_norms) {
    const float* sq = reinterpret_cast<const float*>(host + real * (sizeof(float) + sizeof(int)));
    double s0 = 0.0, s1 = 0.0;
    for (int b = 0; b < nb; ++b) {
      s0 += double(sq[b]);
      s1 += double(sq[nb + b]);
    }
    out.norms[0] = std::sqrt(s0);
    out.norms[1] = std::sqrt(s1);
  }
  return out;
}

// ------------------------------------------------------------------------- field ops (a8-a13) ---

MetricPair<HField> HipBackend::cc_pair_deferred(const HImage& fixed, const HImage& moving, size_t radius) {
  if (fixed.grid != moving.grid) invalid("hip backend", "cc: the images are not on one grid");
  if (radius == 0 || radius > 8) invalid("hip backend", "cc radius " + std::to_string(radius) + " outside 1..=8");
  const Grid& g = fixed.grid;
  const size_t n = g.num_voxels();
  MetricPair<HField> out;
  out.fields[0] = HField{g, DevBuf(ctx_.get(), 3 * n)};
  out.fields[1] = HField{g, DevBuf(ctx_.get(), 3 * n)};
  CcArgs a{};
  a.fixed = fixed.buf.get();
  a.moving = moving.buf.get();
  a.valid = nullptr;
  a.d = dims_of(g);
  a.radius = int(radius);
  a.tile_z = cc_tile_for(a.d, int(ctx_->tuning().get("cc.tile_z", n, cc_tile_z)),
                         int(ctx_->tuning().get("cc.min_blocks", n, 512)));
  a.scale_f = cc_split_scale(fixed.bound);
  a.scale_m = cc_split_scale(moving.bound);
  a.step_f = 1.0f / a.scale_f;
  a.step_m = 1.0f / a.scale_m;
  a.probes = probes_of(g);
  a.d_fixed = out.fields[0].buf.get();
  a.d_moving = out.fields[1].buf.get();
  const dim3 gr = cc_grid(a.d, a.tile_z);
  const size_t blocks = size_t(gr.x) * gr.y * gr.z;
  const size_t real = blocks;
  char* host = nullptr;
  DevBuf partials;
  if (blocks <= kCcSlotBlocks) {
    // the kernel writes its block partials straight into a slot of a ring in coherent host memory,
    // read after the settle (the next evaluation may run before it, in the next slot)
    void* ring_dev = nullptr;
    const size_t bytes = kCcSlotBlocks * (sizeof(float) + sizeof(int));
    char* ring = static_cast<char*>(ctx_->mapped(4, kCcSlots * bytes, &ring_dev));
    const size_t slot = ctx_->next_tag() % kCcSlots;
    host = ring + slot * bytes;
    a.partial_cc = reinterpret_cast<float*>(static_cast<char*>(ring_dev) + slot * bytes);
    a.partial_count = reinterpret_cast<int*>(static_cast<char*>(ring_dev) + slot * bytes + real * sizeof(float));
    Timed t(*ctx_, "k_cc2", 32.0 * double(n));
    launch_cc_pair(a, ctx_->stream());
  } else {
    // the partials live in a buffer of this call's own, copied to staging memory
    partials = DevBuf(ctx_.get(), 2 * blocks);
    a.partial_cc = partials.get();
    a.partial_count = reinterpret_cast<int*>(partials.get() + blocks);
    {
      Timed t(*ctx_, "k_cc2", 32.0 * double(n));
      launch_cc_pair(a, ctx_->stream());
    }
    host = static_cast<char*>(ctx_->stage(real * (sizeof(float) + sizeof(int))));
    check(hipMemcpyAsync(host, a.partial_cc, real * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "cc partials");
    check(hipMemcpyAsync(host + real * sizeof(float), a.partial_count, real * sizeof(int), hipMemcpyDeviceToHost,
                         ctx_->stream()),
          "cc counts");
  }
  out.later = Pending<std::array<double, 2>>(std::function<std::array<double, 2>()>([host, real] {
    const float* pcc = reinterpret_cast<const float*>(host);
    const int* pcount = reinterpret_cast<const int*>(host + real * sizeof(float));
    double total = 0.0;
    long long count = 0;
    for (size_t b = 0; b < real; ++b) {
      total += double(pcc[b]);
      count += pcount[b];
    }
    if (count == 0) invalid("cc metric", "no virtual voxel maps inside both image buffers");
    const double v = -total / double(count);
    return std::array<double, 2>{v, v};
  }));
  return out;
}

std::pair<Pending<std::pair<double, double>>, DevBuf> HipBackend::step_factors(const HField& a, const HField& b, double step) {
  if (a.grid != b.grid) invalid("hip backend", "update pair: grids differ");
  const long long n = long(a.grid.num_voxels());
  const int nb = norm_blocks(n);
  float* pm = static_cast<float*>(ctx_->scratch(2, 2 * nb * sizeof(float)));
  float* ps = static_cast<float*>(ctx_->scratch(3, 2 * nb * sizeof(float)));
  DevBuf factors(ctx_.get(), 2);
  hipStream_t s = ctx_->stream();
  {
    Timed t(*ctx_, "k_field_norms", 24.0 * double(n));
    launch_field_norms(a.buf.get(), b.buf.get(), n, inv_spacing(a.grid), pm, ps, nb, s);
  }
  // the maxima go to a slot of a ring in coherent host memory, read after the settle
  void* ring_dev = nullptr;
  auto* ring = static_cast<float*>(ctx_->mapped(3, kMaximaSlots * 2 * sizeof(float), &ring_dev));
  const size_t slot = ctx_->next_tag() % kMaximaSlots;
  launch_step_factors(pm, nb, 2, step, factors.get(), static_cast<float*>(ring_dev) + 2 * slot, s);
  volatile float* host = ring + 2 * slot;
  return {Pending<std::pair<double, double>>(
              std::function<std::pair<double, double>()>([host] { return std::make_pair(double(host[0]), double(host[1])); })),
          std::move(factors)};
}

Pending<std::pair<double, double>> HipBackend::scale_to_step(HField& a, HField& b, double step) {
  auto r = step_factors(a, b, step);
  const long long n = long(a.grid.num_voxels());
  {
    Timed t(*ctx_, "k_scale_by", 48.0 * double(n));
    launch_scale_by2(a.buf.get(), b.buf.get(), 3 * n, r.second.get(), ctx_->stream());
  }
  return std::move(r.first);
}

HImage HipBackend::smooth_image(const HImage& src, double sigma, SigmaUnits units) {
  const Vec3& sp = src.grid.spacing;
  if (sigma <= 0.0) return smooth_image_mm(src, {0.0, 0.0, 0.0});
  const Vec3 mm = units == SigmaUnits::Voxels ? Vec3{sigma * sp[0], sigma * sp[1], sigma * sp[2]} : Vec3{sigma, sigma, sigma};
  return smooth_image_mm(src, mm);
}

HImage HipBackend::smooth_image_mm(const HImage& src, const Vec3& sigma_mm) {
  const Grid& g = src.grid;
  const size_t n = g.num_voxels();
  HImage out{g, DevBuf(ctx_.get(), n), src.bound};
  hipStream_t s = ctx_->stream();
  if (!(sigma_mm[0] > 0.0 || sigma_mm[1] > 0.0 || sigma_mm[2] > 0.0)) {
    check(hipMemcpyAsync(out.buf.get(), src.buf.get(), n * sizeof(float), hipMemcpyDeviceToDevice, s), "copy image");
    return out;
  }
  for (int axis = 0; axis < 3; ++axis) {
    if (g.size[axis] < 4) invalid("smoothing", "every axis needs at least 4 voxels");
    if (!(sigma_mm[axis] > 0.0)) invalid("smoothing", "sigma must be greater than zero");
  }
  const Dims d = dims_of(g);
  DevBuf tmp(ctx_.get(), n);
  // z into out, x into tmp, y into out (pyramid.rs:429-437 order)
  const float* in = src.buf.get();
  float* bufs[3] = {out.buf.get(), tmp.get(), out.buf.get()};
  const int axes[3] = {2, 0, 1};
  for (int q = 0; q < 3; ++q) {
    const Deriche k = Deriche::zero_order(sigma_mm[axes[q]], g.spacing[axes[q]]);
    DericheCoefficients c{};
    for (int t = 0; t < 4; ++t) {
      c.n[t] = k.n[t];
      c.d[t] = k.d[t];
      c.m[t] = k.m[t];
      c.bn[t] = k.bn[t];
      c.bm[t] = k.bm[t];
    }
    // one read and one write per voxel (the anti-causal pass re-reads its line from cache)
    Timed t(*ctx_, "k_deriche_axis", 8.0 * double(n));
    launch_deriche_axis(in, bufs[q], d, axes[q], c, s, ctx_->tuning().get("deriche.lds", uint64_t(n), 1) != 0);
    in = bufs[q];
  }
  return out;
}

void HipBackend::sample_values(const HImage& img, const HostVec<Vec3>& points, HostVec<double>& nearest,
                               HostVec<double>& linear) {
  const size_t m = points.size();
  nearest.resize(m);  // every element is written below
  linear.resize(m);
  if (m == 0) return;
  const Grid& g = img.grid;
  hipStream_t s = ctx_->stream();
  auto* stage = static_cast<float*>(ctx_->pinned(4 * m * sizeof(float)));
  parallel_for(m, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const Vec3 c = g.point_to_index(points[i]);
      stage[4 * i] = float(c[0]);
      stage[4 * i + 1] = float(c[1]);
      stage[4 * i + 2] = float(c[2]);
      stage[4 * i + 3] = 0.0f;
    }
  });
  DevBuf idx(ctx_.get(), 4 * m), values(ctx_.get(), 2 * m);
  check(hipMemcpyAsync(idx.get(), stage, 4 * m * sizeof(float), hipMemcpyHostToDevice, s), "sample indices");
  {
    Timed t(*ctx_, "k_sample_values", 16.0 * double(m) + 8.0 * double(m));
    launch_sample_values(img.buf.get(), dims_of(g), reinterpret_cast<const float4*>(idx.get()), (long long)m,
                         values.get(), values.get() + m, s);
  }
  // the stage is reused for the values: stream order puts this copy after the index upload
  check(hipMemcpyAsync(stage, values.get(), 2 * m * sizeof(float), hipMemcpyDeviceToHost, s), "sample values");
  ctx_->sync();
  parallel_for(m, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      nearest[i] = double(stage[i]);
      linear[i] = double(stage[m + i]);
    }
  });
}

// preprocess.rs:159-232 on an image in HBM: the finite range and the non-finite count, the 256-bin
// histogram (bins by f32 thresholds equal to the f64 bins), the quantile window on the host (a few
// hundred f64 operations) and the window map on the device.
HImage HipBackend::preprocess_image(const HImage& raw, float lower, float upper, double out_min, double out_max) {
  const Grid& g = raw.grid;
  const long long n = (long long)g.num_voxels();
  const int nb = norm_blocks(n);
  hipStream_t s = ctx_->stream();
  auto* part = static_cast<float*>(ctx_->scratch(2, 3 * size_t(nb) * sizeof(float)));
  {
    Timed t(*ctx_, "k_range_and_bad", 4.0 * double(n));
    launch_range_and_bad(raw.buf.get(), n, part, part + nb, reinterpret_cast<unsigned int*>(part + 2 * nb), nb, s);
  }
  auto* host = static_cast<float*>(ctx_->pinned(3 * size_t(nb) * sizeof(float)));
  check(hipMemcpyAsync(host, part, 3 * size_t(nb) * sizeof(float), hipMemcpyDeviceToHost, s), "range partials");
  ctx_->sync();
  float lo = INFINITY, hi = -INFINITY;
  uint64_t bad = 0;
  for (int b = 0; b < nb; ++b) {
    lo = std::min(lo, host[b]);
    hi = std::max(hi, host[nb + b]);
    unsigned int c;
    std::memcpy(&c, &host[2 * nb + b], sizeof c);
    bad += c;
  }
  if (bad) invalid("register", "an input image has non-finite (NaN or infinite) voxels");
  if (!(lo <= hi)) invalid("preprocess", "image has no finite values");
  const IntensityHistogram edges(double(lo), double(hi), std::vector<uint64_t>(kIntensityBins, 0));
  const std::vector<float> th = edges.f32_thresholds();
  auto* th_host = static_cast<float*>(ctx_->stage(th.size() * sizeof(float)));
  std::memcpy(th_host, th.data(), th.size() * sizeof(float));
  auto* th_dev = static_cast<float*>(ctx_->scratch(3, th.size() * sizeof(float)));
  check(hipMemcpyAsync(th_dev, th_host, th.size() * sizeof(float), hipMemcpyHostToDevice, s), "bin thresholds");
  auto* slots = static_cast<unsigned long long*>(ctx_->scratch(6, size_t(nb) * kIntensityBins * sizeof(unsigned long long)));
  {
    Timed t(*ctx_, "k_intensity_hist", 4.0 * double(n));
    launch_intensity_hist(raw.buf.get(), n, th_dev, slots, nb, s);
  }
  void* dev_counts = nullptr;
  auto* counts = static_cast<unsigned long long*>(ctx_->mapped(0, kIntensityBins * sizeof(unsigned long long), &dev_counts));
  const uint32_t tag = ctx_->next_tag();
  {
    Timed t(*ctx_, "k_mi_sum_slots", 8.0 * double(kIntensityBins) * double(nb + 1));
    launch_mi_sum_slots(slots, nb, kIntensityBins, static_cast<unsigned long long*>(dev_counts), ctx_->done_words_device(),
                        tag, s);
  }
  ctx_->wait_done(mi_sum_slots_blocks(kIntensityBins), tag);
  ctx_->settle();
  const IntensityHistogram h(double(lo), double(hi), std::vector<uint64_t>(counts, counts + kIntensityBins));
  const IntensityWindow w = intensity_window(h, lower, upper, out_min, out_max);
  const float sh = float(w.scale), th_ = float(w.shift);
  const float sl = float(w.scale - double(sh)), tl = float(w.shift - double(th_));
  HImage out{g, DevBuf(ctx_.get(), size_t(n)), float(std::max(std::fabs(out_min), std::fabs(out_max)))};
  {
    Timed t(*ctx_, "k_window", 8.0 * double(n));
    launch_window(raw.buf.get(), out.buf.get(), n, float(w.lo), float(w.hi), sh, sl, th_, tl, float(out_min), float(out_max), s);
  }
  return out;
}

// init.rs:47-86 (ITK ImageMomentsCalculator) on an image in HBM: compensated f32 block sums of
// v, v i, v j, v k over the non-zero voxels, added in f64 on the host in block order.
Vec3 HipBackend::centre_of_gravity(const HImage& img) {
  const Grid& g = img.grid;
  const Dims d = dims_of(g);
  const dim3 gr = moments_grid(d);
  const size_t blocks = size_t(gr.x) * gr.y * gr.z;
  auto* part = static_cast<float*>(ctx_->scratch(2, 4 * blocks * sizeof(float)));
  hipStream_t s = ctx_->stream();
  {
    Timed t(*ctx_, "k_moments", 4.0 * double(g.num_voxels()));
    launch_moments(img.buf.get(), d, part, s);
  }
  auto* host = static_cast<float*>(ctx_->pinned(4 * blocks * sizeof(float)));
  check(hipMemcpyAsync(host, part, 4 * blocks * sizeof(float), hipMemcpyDeviceToHost, s), "moment partials");
  ctx_->sync();
  double sum[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t b = 0; b < blocks; ++b)
    for (int q = 0; q < 4; ++q) sum[q] += double(host[4 * b + q]);
  if (sum[0] == 0.0) invalid("initial transform", "total mass of the image is zero");
  return g.index_to_point({sum[1] / sum[0], sum[2] / sum[0], sum[3] / sum[0]});
}

bool HipBackend::finite_range(const HImage& img, double& lo, double& hi) {
  const long long n = (long long)img.grid.num_voxels();
  const int nb = norm_blocks(n);
  float* p = static_cast<float*>(ctx_->scratch(2, 2 * nb * sizeof(float)));
  {
    Timed t(*ctx_, "k_finite_range", 4.0 * double(n));
    launch_finite_range(img.buf.get(), n, p, p + nb, nb, ctx_->stream());
  }
  float* host = static_cast<float*>(ctx_->pinned(2 * nb * sizeof(float)));
  check(hipMemcpyAsync(host, p, 2 * nb * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "range partials");
  ctx_->sync();
  float l = INFINITY, h = -INFINITY;
  for (int b = 0; b < nb; ++b) {
    l = std::min(l, host[b]);
    h = std::max(h, host[nb + b]);
  }
  if (!(l <= h)) return false;
  lo = double(l);
  hi = double(h);
  return true;
}

std::pair<HField, HField> HipBackend::smooth_field_pair(HField a, HField b, double variance) {
  if (variance <= 0.0) return {std::move(a), std::move(b)};
  const Grid& g = a.grid;
  if (b.grid != g) return {smooth_field(std::move(a), variance), smooth_field(std::move(b), variance)};
  const Dims d = dims_of(g);
  const size_t n = g.num_voxels();
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
  GaussTaps3 t{};
  int* radii[3] = {&t.rx, &t.ry, &t.rz};
  float* dst[3] = {t.x, t.y, t.z};
  for (int q = 0; q < 3; ++q) {
    const std::vector<double>& taps = ctx_->host_taps(variance, g.size[q]);
    *radii[q] = int(taps.size() / 2);
    if (taps.size() > 17) return {smooth_field(std::move(a), variance), smooth_field(std::move(b), variance)};
    for (size_t k = 0; k < taps.size(); ++k) dst[q][k] = float(taps[k]);
  }
  if (!gauss3_fits(d, t)) return {smooth_field(std::move(a), variance), smooth_field(std::move(b), variance)};
  HField out0{g, DevBuf(ctx_.get(), 3 * n)}, out1{g, DevBuf(ctx_.get(), 3 * n)};
  if (t.rx == t.ry && t.ry == t.rz && ctx_->tuning().get("gauss.one_pass", uint64_t(d.n()), gauss3_pays(d) ? 1 : 0)) {
    // both fields read once (twice with a blend), both results written
    Timed tm(*ctx_, "k_gauss3", 48.0 * double(n) + (w2 != 0.0 ? 24.0 * double(n) : 0.0));
    if (!launch_gauss3(a.buf.get(), out0.buf.get(), d, t, float(w1), float(w2), ctx_->stream(), b.buf.get(), out1.buf.get()))
      invalid("smoothing", "no one-pass kernel for radius " + std::to_string(t.rz));
    return {std::move(out0), std::move(out1)};
  }
  DevBuf tmp0(ctx_.get(), 3 * n), tmp1(ctx_.get(), 3 * n);
  {
    Timed tm(*ctx_, "k_gauss_xy", 48.0 * double(n));
    launch_gauss_xy(a.buf.get(), tmp0.get(), d, t, ctx_->stream(), b.buf.get(), tmp1.get());
  }
  {
    Timed tm(*ctx_, "k_gauss_z", 48.0 * double(n) + (w2 != 0.0 ? 24.0 * double(n) : 0.0));
    launch_gauss_z(tmp0.get(), a.buf.get(), out0.buf.get(), d, t, float(w1), float(w2), ctx_->stream(), tmp1.get(),
                   b.buf.get(), out1.buf.get());
  }
  return {std::move(out0), std::move(out1)};
}

HField HipBackend::smooth_field(HField field, double variance) {
  if (variance <= 0.0) return field;
  const Grid& g = field.grid;
  const Dims d = dims_of(g);
  const size_t n = g.num_voxels();
  const double w1 = variance < 0.5 ? 1.0 - variance / 0.5 : 1.0;
  const double w2 = 1.0 - w1;
  {
    GaussTaps3 t{};
    int* radii[3] = {&t.rx, &t.ry, &t.rz};
    float* dst[3] = {t.x, t.y, t.z};
    bool fits = true;
    for (int a = 0; a < 3; ++a) {
      const std::vector<double>& taps = ctx_->host_taps(variance, g.size[a]);
      *radii[a] = int(taps.size() / 2);
      if (taps.size() > 17) {
        fits = false;
        break;
      }
      for (size_t q = 0; q < taps.size(); ++q) dst[a][q] = float(taps[q]);
    }
    if (fits && gauss3_fits(d, t)) {
      HField out{g, DevBuf(ctx_.get(), 3 * n)};
      if (t.rx == t.ry && t.ry == t.rz && ctx_->tuning().get("gauss.one_pass", uint64_t(d.n()), gauss3_pays(d) ? 1 : 0)) {
        // u read once (twice with a blend), the result written
        Timed tm(*ctx_, "k_gauss3", 24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0));
        if (!launch_gauss3(field.buf.get(), out.buf.get(), d, t, float(w1), float(w2), ctx_->stream()))
          invalid("smoothing", "no one-pass kernel for radius " + std::to_string(t.rz));
        return out;
      }
      DevBuf tmp(ctx_.get(), 3 * n);
      {
        // read u, write the x/y-smoothed intermediate
        Timed tm(*ctx_, "k_gauss_xy", 24.0 * double(n));
        launch_gauss_xy(field.buf.get(), tmp.get(), d, t, ctx_->stream());
      }
      {
        // read the intermediate (and u for a blend), write the result
        Timed tm(*ctx_, "k_gauss_z", 24.0 * double(n) + (w2 != 0.0 ? 12.0 * double(n) : 0.0));
        launch_gauss_z(tmp.get(), field.buf.get(), out.buf.get(), d, t, float(w1), float(w2), ctx_->stream());
      }
      return out;
    }
  }
  DevBuf t1(ctx_.get(), 3 * n), t2(ctx_.get(), 3 * n);
  int r = 0;
  const double nb = double(n);
  const float* tx = ctx_->taps(variance, g.size[0], r);
  {
    Timed t(*ctx_, "k_gauss_axis", 24.0 * nb);
    launch_gauss_axis(field.buf.get(), t1.get(), d, 0, tx, r, nullptr, 1.0f, 0.0f, 0, ctx_->stream());
  }
  const float* ty = ctx_->taps(variance, g.size[1], r);
  {
    Timed t(*ctx_, "k_gauss_axis", 24.0 * nb);
    launch_gauss_axis(t1.get(), t2.get(), d, 1, ty, r, nullptr, 1.0f, 0.0f, 0, ctx_->stream());
  }
  const float* tz = ctx_->taps(variance, g.size[2], r);
  {
    Timed t(*ctx_, "k_gauss_axis", 36.0 * nb);
    launch_gauss_axis(t2.get(), t1.get(), d, 2, tz, r, field.buf.get(), float(w1), float(w2), 1, ctx_->stream());
  }
  return HField{g, std::move(t1)};
}

std::pair<double, double> HipBackend::max_scaled_norm_pair(const HField& a, const HField* b) {
  if (b && b->grid.num_voxels() != a.grid.num_voxels()) invalid("hip backend", "norm pair: grids differ");
  const long long n = long(a.grid.num_voxels());
  const int nb = norm_blocks(n);
  float* pm = static_cast<float*>(ctx_->scratch(2, 2 * nb * sizeof(float)));
  float* ps = static_cast<float*>(ctx_->scratch(3, 2 * nb * sizeof(float)));
  double result[2] = {0.0, 0.0};
  // the two fields may live on different grids only in spacing-free size; run each on its own
  const HField* fields[2] = {&a, b};
  for (int q = 0; q < (b ? 2 : 1); ++q) {
    {
      Timed t(*ctx_, "k_field_norms", 12.0 * double(n));
      launch_field_norms(fields[q]->buf.get(), nullptr, n, inv_spacing(fields[q]->grid), pm, ps, nb, ctx_->stream());
    }
    float* host = static_cast<float*>(ctx_->pinned(nb * sizeof(float)));
    check(hipMemcpyAsync(host, pm, nb * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "norm partials");
    ctx_->sync();
    double m = 0.0;
    for (int i = 0; i < nb; ++i) {
      const double v = host[i];
      if (std::isnan(v) || std::isnan(m)) m = NAN;
      else m = std::max(m, v);
    }
    result[q] = m;
  }
  return {result[0], result[1]};
}

double HipBackend::l2_norm(const HField& f) {
  const long long n = long(f.grid.num_voxels());
  const int nb = norm_blocks(n);
  float* pm = static_cast<float*>(ctx_->scratch(2, 2 * nb * sizeof(float)));
  float* ps = static_cast<float*>(ctx_->scratch(3, 2 * nb * sizeof(float)));
  {
    Timed t(*ctx_, "k_field_norms", 12.0 * double(n));
    launch_field_norms(f.buf.get(), nullptr, n, inv_spacing(f.grid), pm, ps, nb, ctx_->stream());
  }
  float* host = static_cast<float*>(ctx_->pinned(nb * sizeof(float)));
  check(hipMemcpyAsync(host, ps, nb * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream()), "l2 partials");
  ctx_->sync();
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += double(host[i]);
  return std::sqrt(s);
}

void HipBackend::scale(HField& f, double s) {
  Timed t(*ctx_, "k_scale", 24.0 * double(f.grid.num_voxels()));
  launch_scale(f.buf.get(), long(3 * f.grid.num_voxels()), float(s), ctx_->stream());
}

void HipBackend::axpy(HField& acc, double a, const HField& x) {
  if (acc.grid != x.grid) invalid("field axpy", "grids differ");
  Timed t(*ctx_, "k_axpy", 36.0 * double(acc.grid.num_voxels()));
  launch_axpy(acc.buf.get(), x.buf.get(), long(3 * acc.grid.num_voxels()), float(a), ctx_->stream());
}

HField HipBackend::compose(const HField& displacement, const HField& warping) {
  if (displacement.grid != warping.grid) invalid("compose", "displacement and warping grids differ");
  const Grid& g = warping.grid;
  HField out{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  float l[9];
  put(l, g.index_from_offset);
  Timed t(*ctx_, "k_compose", 36.0 * double(g.num_voxels()));
  launch_compose(displacement.buf.get(), warping.buf.get(), dims_of(g), l, out.buf.get(), ctx_->stream());
  return out;
}

std::pair<HField, InvertStats> HipBackend::invert(const HField& field, HField initial, const InvertParams& p) {
  auto r = invert_deferred(field, std::move(initial), p);
  ctx_->settle();
  return {std::move(r.first), r.second.get()};
}

std::pair<HField, HField> HipBackend::compose_pair(const HField& u0, const HField& w0, const HField& u1, const HField& w1) {
  if (u0.grid != w0.grid || u1.grid != w1.grid) invalid("compose", "displacement and warping grids differ");
  if (w0.grid != w1.grid) return {compose(u0, w0), compose(u1, w1)};
  const Grid& g = w0.grid;
  HField out0{g, DevBuf(ctx_.get(), 3 * g.num_voxels())}, out1{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  float l[9];
  put(l, g.index_from_offset);
  Timed t(*ctx_, "k_compose", 72.0 * double(g.num_voxels()));
  launch_compose2(u0.buf.get(), w0.buf.get(), out0.buf.get(), u1.buf.get(), w1.buf.get(), out1.buf.get(), dims_of(g), l,
                  nullptr, ctx_->stream());
  return {std::move(out0), std::move(out1)};
}

std::pair<HField, HField> HipBackend::compose_pair_scaled(const HField& u0, const HField& w0, const HField& u1, const HField& w1,
                                                          const DevBuf& factors) {
  if (u0.grid != w0.grid || u1.grid != w1.grid || w0.grid != w1.grid) invalid("compose", "the pair's grids differ");
  const Grid& g = w0.grid;
  HField out0{g, DevBuf(ctx_.get(), 3 * g.num_voxels())}, out1{g, DevBuf(ctx_.get(), 3 * g.num_voxels())};
  float l[9];
  put(l, g.index_from_offset);
  Timed t(*ctx_, "k_compose", 72.0 * double(g.num_voxels()));
  launch_compose2(u0.buf.get(), w0.buf.get(), out0.buf.get(), u1.buf.get(), w1.buf.get(), out1.buf.get(), dims_of(g), l,
                  factors.get(), ctx_->stream());
  return {std::move(out0), std::move(out1)};
}

std::pair<HField, Pending<InvertStats>> HipBackend::invert_deferred(const HField& field, HField initial,
                                                                    const InvertParams& p) {
  if (initial.grid != field.grid) invalid("invert", "initial estimate grid differs");
  const Grid& g = field.grid;
  const Dims d = dims_of(g);
  // the kernels split voxel indices with 32-bit divisions
  if (d.n() >= (1ll << 31)) invalid("invert", "grids of 2^31 voxels or more are not supported");
  float l[9];
  put(l, g.index_from_offset);
  const float3 isp = inv_spacing(g);
  const int nb = inv_blocks(d, int(ctx_->tuning().get("inv.max_blocks", uint64_t(d.n()), 2048)));
  auto* state = static_cast<InvertState*>(ctx_->scratch(4, sizeof(InvertState)));
  auto* partials = static_cast<float*>(ctx_->scratch(5, 4 * size_t(nb) * sizeof(float)));
  hipStream_t s = ctx_->stream();
  const int max_it = int(p.max_iterations);
  // The iterations' launches are enqueued one by one; once the device has reported (through the
  // mailbox) that this inversion stopped, the remaining ones, which would return at once, are not.
  volatile uint32_t* mail = ctx_->mailbox_host();
  uint32_t* mail_dev = ctx_->mailbox_device();
  const uint32_t tag = ctx_->next_tag();
  DevBuf residual(ctx_.get(), 3 * g.num_voxels());
  // the inversion's report: a slot of a ring in coherent host memory, read after the settle
  constexpr size_t kReports = 64;
  void* report_dev = nullptr;
  auto* reports = static_cast<InvertReport*>(ctx_->mapped(2, kReports * sizeof(InvertReport), &report_dev));
  const size_t slot = tag % kReports;
  InvertReport* report = static_cast<InvertReport*>(report_dev) + slot;
  volatile InvertReport* host = reports + slot;
  {
    Timed t(*ctx_, "k_inv_begin", 36.0 * double(d.n()));
    launch_inv_begin(field.buf.get(), initial.buf.get(), residual.get(), d, l, isp, state, report, partials, nb, max_it, s);
  }
  // At most kAhead launches are queued beyond the one the device has started: the host waits on
  // the progress word (a hint; a bounded wait) so that it does not queue launches past the stop.
  constexpr int kAhead = 1;
  const auto started = [&]() -> int {
    const uint32_t w = mail[1];
    return (w >> 8) == (tag & 0xffffffu) ? int(w & 0xffu) : 0;
  };
  for (int step = 0; step < max_it; ++step) {
    if (step > 0 && mail[0] == tag) break;
    if (step >= kAhead) {
      const auto t0 = std::chrono::steady_clock::now();
      for (unsigned spins = 1; started() < step - kAhead + 1 && mail[0] != tag; ++spins)
        if ((spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
      if (mail[0] == tag) break;
    }
    Timed t(*ctx_, "k_inv_iter", 0.0);
    launch_inv_iter(field.buf.get(), initial.buf.get(), residual.get(), d, l, isp, state, report, partials, nb, step, max_it,
                    p.max_error_tolerance, p.mean_error_tolerance, p.enforce_boundary_condition ? 1 : 0, mail_dev, tag, s);
  }
  DeviceContext* ctx = ctx_.get();
  const double voxels = double(d.n());
  const bool profiling = ctx_->profiling();
  return {std::move(initial), Pending<InvertStats>(std::function<InvertStats()>([host, ctx, voxels, profiling] {
            InvertStats st;
            st.iterations = uint32_t(host->executed);
            // an iteration that another follows reads psi, the stored residual and the field and
            // writes psi and the new residual (60 B per voxel); the last one reads psi and the
            // residual and writes psi (36 B)
            if (profiling && host->executed > 0)
              ctx->add_bytes("k_inv_iter", voxels * (60.0 * double(host->executed - 1) + 36.0));
            st.mean_error_norm = host->executed ? double(host->mean_error) : 0.0;
            st.max_error_norm = host->executed ? double(host->max_error) : 0.0;
            return st;
          }))};
}

HField HipBackend::resample_field(const HField& field, const Grid& grid) {
  if (field.grid == grid) return copy_field(field);
  const Grid& sg = field.grid;
  ResampleArgs a{};
  a.src = field.buf.get();
  a.sd = dims_of(sg);
  a.od = dims_of(grid);
  put(a.to_src.m, mat3_mul(sg.index_from_offset, scaled_direction(grid)));
  put(a.to_src.c0, mat3_mul_vec(sg.index_from_offset, vsub(grid.origin, sg.origin)));
  HField out{grid, DevBuf(ctx_.get(), 3 * grid.num_voxels())};
  a.out = out.buf.get();
  Timed t(*ctx_, "k_resample", 12.0 * double(grid.num_voxels()) + 12.0 * std::min(double(grid.num_voxels()), double(sg.num_voxels())));
  launch_resample(a, ctx_->stream());
  return out;
}

}  // namespace fw::hip
