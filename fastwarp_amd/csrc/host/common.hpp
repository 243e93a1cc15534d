// Host-side foundations shared by the registration drivers: 3-vectors and 3x3 matrices in f64
// (row-major, m[row][col]), the error type every layer throws, and a deterministic parallel loop.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <functional>
#include <mutex>
#include <initializer_list>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace fw {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

// Large host blocks kept for reuse: a registration allocates the same large buffers (sample
// points, sample values, preprocessed volumes) every time, and a block that has been touched
// once costs no page faults when it is handed out again. A request takes the smallest cached
// block that holds it, if that block is at most a quarter larger; blocks beyond the cache's
// limit (FW_HOST_CACHE_MB, default 2048 MiB) go back to the system, oldest first.
class HostBlockCache {
 public:
  static constexpr size_t kMinBytes = size_t(1) << 20;
  static HostBlockCache& get() {
    static HostBlockCache* cache = new HostBlockCache();  // never destroyed: blocks live until exit
    return *cache;
  }
  void* take(size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      size_t best = free_.size();
      for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].bytes >= bytes && free_[i].bytes <= bytes + bytes / 4 &&
            (best == free_.size() || free_[i].bytes < free_[best].bytes))
          best = i;
      if (best != free_.size()) {
        const Block b = free_[best];
        free_.erase(free_.begin() + std::ptrdiff_t(best));
        cached_ -= b.bytes;
        live_.push_back(b);
        return b.p;
      }
    }
    void* p = ::operator new(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    live_.push_back({p, bytes});
    return p;
  }
  void give(void* p) {
    std::vector<void*> release;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = std::find_if(live_.begin(), live_.end(), [&](const Block& b) { return b.p == p; });
      if (it == live_.end()) {
        release.push_back(p);
      } else {
        free_.push_back(*it);
        cached_ += it->bytes;
        live_.erase(it);
        while (cached_ > limit_ && !free_.empty()) {
          release.push_back(free_.front().p);
          cached_ -= free_.front().bytes;
          free_.erase(free_.begin());
        }
      }
    }
    for (void* q : release) ::operator delete(q);
  }

 private:
  HostBlockCache() {
    if (const char* s = std::getenv("FW_HOST_CACHE_MB")) {
      const long v = std::strtol(s, nullptr, 10);
      if (v >= 0) limit_ = size_t(v) << 20;
    }
  }
  struct Block {
    void* p;
    size_t bytes;
  };
  std::mutex mu_;
  std::vector<Block> free_, live_;
  size_t cached_ = 0, limit_ = size_t(2048) << 20;
};

// An allocator whose value-less construct leaves the element default-initialised (indeterminate
// for arithmetic types and arrays of them), so that resizing a large host buffer does not write it
// serially first: the parallel loop that fills it is the first to touch its pages. Blocks of a
// MiB and more come from the HostBlockCache.
template <typename T>
struct DefaultInit : std::allocator<T> {
  using value_type = T;
  template <typename U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() = default;
  template <typename U>
  DefaultInit(const DefaultInit<U>&) noexcept {}
  T* allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes >= HostBlockCache::kMinBytes) return static_cast<T*>(HostBlockCache::get().take(bytes));
    return std::allocator<T>::allocate(n);
  }
  void deallocate(T* p, size_t n) {
    if (n * sizeof(T) >= HostBlockCache::kMinBytes) HostBlockCache::get().give(p);
    else std::allocator<T>::deallocate(p, n);
  }
  template <typename U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void*>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <typename T>
using HostVec = std::vector<T, DefaultInit<T>>;

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

// device.rs:919-928: the FASTWARP_GPU_WAIT_TIMEOUT_S value if it parses as positive finite
// seconds, else 300.
inline double wait_timeout_from(const char* text) {
  constexpr double kDefault = 300.0;
  if (!text) return kDefault;
  char* end = nullptr;
  const double v = std::strtod(text, &end);
  while (end && (*end == ' ' || *end == '\t' || *end == '\n')) ++end;
  if (end == text || (end && *end != '\0') || !std::isfinite(v) || !(v > 0.0)) return kDefault;
  return v;
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
//
// The loops run on one process-wide pool of host_threads() - 1 parked workers (a new thread per
// loop would cost tens of microseconds per thread). Every loop is a job in the pool's list; the
// calling thread works on its own job while idle workers take chunks of the oldest job with
// chunks left, so loops started on several threads at once (the drivers' main thread and the
// sample prefetch) share the workers instead of each starting a full set of threads. A loop body
// may itself run a loop.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* pool = new HostPool(host_threads());  // never destroyed: its workers stay parked
    return *pool;
  }
  // body(c) for every chunk c < chunks; the first exception a chunk throws is rethrown here.
  void run(size_t chunks, const std::function<void(size_t)>& body) {
    Job job;
    job.body = &body;
    job.chunks = chunks;
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_.push_back(&job);
      generation_.fetch_add(1, std::memory_order_release);
    }
    wake_.notify_all();
    work_on(job);
    // every chunk is claimed now; wait for the workers still running one of them
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_.erase(std::find(jobs_.begin(), jobs_.end(), &job));
    }
    if (!spin_until([&] { return job.users.load(std::memory_order_acquire) == 0; })) {
      std::unique_lock<std::mutex> lk(mu_);
      done_.wait(lk, [&] { return job.users.load(std::memory_order_acquire) == 0; });
    }
    if (job.error) std::rethrow_exception(job.error);
  }

 private:
  struct Job {
    const std::function<void(size_t)>* body = nullptr;
    size_t chunks = 0;
    std::atomic<size_t> next{0};
    std::atomic<unsigned> users{0};  // workers inside work_on (raised under mu_)
    std::exception_ptr error;
  };
  explicit HostPool(unsigned threads) {
    for (unsigned t = 1; t < threads; ++t) workers_.emplace_back([this] { loop(); });
    for (auto& w : workers_) w.detach();
  }
  // Polls `ready` for about 100 microseconds (a parked thread takes longer than that to wake).
  template <typename F>
  static bool spin_until(F ready) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned i = 1;; ++i) {
      if (ready()) return true;
      if ((i & 63) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(100)) return ready();
    }
  }
  void work_on(Job& job) {
    for (;;) {
      const size_t c = job.next.fetch_add(1);
      if (c >= job.chunks) return;
      try {
        (*job.body)(c);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!job.error) job.error = std::current_exception();
        job.next.store(job.chunks);
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      Job* job = nullptr;
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (Job* j : jobs_)
          if (j->next.load() < j->chunks) {
            job = j;
            break;
          }
        if (job) job->users.fetch_add(1, std::memory_order_relaxed);
        else seen = generation_.load(std::memory_order_acquire);
      }
      if (job) {
        work_on(*job);
        if (job->users.fetch_sub(1, std::memory_order_acq_rel) == 1) {
          std::lock_guard<std::mutex> lk(mu_);
          done_.notify_all();
        }
        continue;
      }
      if (!spin_until([&] { return generation_.load(std::memory_order_acquire) != seen; })) {
        std::unique_lock<std::mutex> lk(mu_);
        wake_.wait(lk, [&] { return generation_.load(std::memory_order_acquire) != seen; });
      }
    }
  }
  std::mutex mu_;
  std::condition_variable wake_, done_;
  std::vector<std::thread> workers_;
  std::vector<Job*> jobs_;
  std::atomic<uint64_t> generation_{0};
};

inline void parallel_chunks(size_t n, size_t grain, const std::function<void(size_t, size_t, size_t)>& body) {
  if (n == 0) return;
  grain = std::max<size_t>(grain, 1);
  const size_t chunks = (n + grain - 1) / grain;
  auto chunk = [&](size_t c) { body(c, c * grain, std::min(n, (c + 1) * grain)); };
  if (chunks == 1 || host_threads() <= 1) {
    for (size_t c = 0; c < chunks; ++c) chunk(c);
    return;
  }
  HostPool::get().run(chunks, chunk);
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
