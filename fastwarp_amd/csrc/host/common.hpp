// Host-side foundations shared by the registration drivers: 3-vectors and 3x3 matrices in f64
// (row-major, m[row][col]), the error type every layer throws, and a deterministic parallel loop.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace fw {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

constexpr Mat3 kIdentity3 = {{{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}};
constexpr double kEps = std::numeric_limits<double>::epsilon();

// The error every operation raises. `what` names the operation or layer, `message` says what was
// wrong; the C ABI reports "what: message" (crates/fastwarp-reg/src/error.rs `RegError::Invalid`).
struct Error : std::runtime_error {
  std::string what_;
  std::string message;
  Error(std::string w, std::string m)
      : std::runtime_error(w + ": " + m), what_(std::move(w)), message(std::move(m)) {}
};

// Raised when the cancel flag of a run is set (crates/fastwarp-reg/src/error.rs `Cancelled`).
struct Cancelled : std::runtime_error {
  Cancelled() : std::runtime_error("the registration was cancelled") {}
};

[[noreturn]] inline void invalid(const std::string& what, const std::string& message) {
  throw Error(what, message);
}

inline Vec3 mat3_mul_vec(const Mat3& m, const Vec3& v) {
  return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
          m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
          m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}

inline Mat3 mat3_mul(const Mat3& a, const Mat3& b) {
  Mat3 out{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
  return out;
}

inline double mat3_det(const Mat3& m) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// Inverse by the adjugate; false when the determinant is zero or not finite.
inline bool mat3_inverse(const Mat3& m, Mat3& out) {
  const double det = mat3_det(m);
  if (det == 0.0 || !std::isfinite(det)) return false;
  const double k = 1.0 / det;
  out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * k;
  out[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * k;
  out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * k;
  out[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * k;
  out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * k;
  out[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * k;
  out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * k;
  out[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * k;
  out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * k;
  return true;
}

inline Vec3 vadd(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Vec3 vsub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double vnorm(const Vec3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Wall-clock spans of host work. While a sink is installed on the calling thread (a device
// context does so while it profiles), every HostTimer reports its name and duration to it.
struct HostSpanSink {
  virtual void host_span(const char* name, double seconds) = 0;
  virtual ~HostSpanSink() = default;
};
inline HostSpanSink*& host_span_sink() {
  thread_local HostSpanSink* sink = nullptr;
  return sink;
}
class HostTimer {
 public:
  explicit HostTimer(const char* name) : name_(name), sink_(host_span_sink()) {
    if (sink_) t0_ = std::chrono::steady_clock::now();
  }
  ~HostTimer() {
    if (sink_) sink_->host_span(name_, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count());
  }
  HostTimer(const HostTimer&) = delete;
  HostTimer& operator=(const HostTimer&) = delete;

 private:
  const char* name_;
  HostSpanSink* sink_;
  std::chrono::steady_clock::time_point t0_;
};

// Number of worker threads of the host loops: FW_THREADS, else OMP_NUM_THREADS, else the hardware count.
inline unsigned host_threads() {
  static unsigned n = [] {
    for (const char* name : {"FW_THREADS", "OMP_NUM_THREADS"}) {
      if (const char* s = std::getenv(name)) {
        long v = std::strtol(s, nullptr, 10);
        if (v > 0) return static_cast<unsigned>(v);
      }
    }
    unsigned h = std::thread::hardware_concurrency();
    return h == 0 ? 1u : h;
  }();
  return n;
}

// Runs body(begin, end) over [0, n) in contiguous chunks on the host threads. Each chunk is a
// fixed function of n and `grain`, so per-chunk results combined in chunk order do not depend on
// the thread count.
inline void parallel_chunks(size_t n, size_t grain, const std::function<void(size_t, size_t, size_t)>& body) {
  if (n == 0) return;
  grain = std::max<size_t>(grain, 1);
  const size_t chunks = (n + grain - 1) / grain;
  const unsigned threads = static_cast<unsigned>(std::min<size_t>(host_threads(), chunks));
  if (threads <= 1) {
    for (size_t c = 0; c < chunks; ++c) body(c, c * grain, std::min(n, (c + 1) * grain));
    return;
  }
  std::atomic<size_t> next{0};
  auto worker = [&] {
    for (;;) {
      size_t c = next.fetch_add(1);
      if (c >= chunks) return;
      body(c, c * grain, std::min(n, (c + 1) * grain));
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(threads - 1);
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
}

// parallel_chunks with a default grain that gives every thread several chunks.
inline void parallel_for(size_t n, const std::function<void(size_t, size_t)>& body) {
  const size_t grain = std::max<size_t>(4096, n / (host_threads() * 8 + 1));
  parallel_chunks(n, grain, [&](size_t, size_t b, size_t e) { body(b, e); });
}

// Sum of `values` in fixed chunks of 4096 combined in order, so the result is independent of the
// thread count (crates/fastwarp-reg/src/field_ops.rs:47-56 `ordered_sum`).
inline double ordered_sum(const std::vector<double>& values) {
  constexpr size_t kChunk = 4096;
  const size_t chunks = (values.size() + kChunk - 1) / kChunk;
  std::vector<double> partial(chunks, 0.0);
  parallel_chunks(values.size(), kChunk, [&](size_t c, size_t b, size_t e) {
    double s = 0.0;
    for (size_t i = b; i < e; ++i) s += values[i];
    partial[c] = s;
  });
  double total = 0.0;
  for (double p : partial) total += p;
  return total;
}

}  // namespace fw
