// The registration server of include/fastwarp_hip.h: continuous batching of whole registrations
// on one device. Each slot is a worker thread with its own HipBackend (its own HIP stream and
// buffer pool), the reference's model of several registrations on one GPU as independent device
// instances (crates/fastwarp-cli/src/engine.rs:1015-1068). Requests are admitted by device-memory
// footprint as slots and memory free up (host/admission.hpp), their kernels overlap on the device
// through the slots' streams, and a request's result does not depend on what runs beside it: a
// slot's work reads and writes only its own buffers, and every reduction has a fixed order.
#include <chrono>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fastwarp_hip.h"
#include "hip/backend.hpp"
#include "host/admission.hpp"
#include "host/capacity.hpp"
#include "plan.hpp"

namespace fw::serve {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  fw_plan plan{};
  Grid fixed_grid, moving_grid;
  std::vector<const float*> fixed, moving;
  double* forward = nullptr;
  double* inverse = nullptr;
  uint64_t footprint = 0;
  volatile int32_t cancel = 0;
  Clock::time_point submitted, started, finished;
  int32_t slot = -1;
  bool done = false;
  int status = FW_OK;
  std::string error;
  fw_result result{};
};

class Server {
 public:
  Server(int device, uint32_t slots, uint64_t budget) : device_(device), queue_(budget) {
    if (slots < 1 || slots > kMaxSlots) invalid("server", "between 1 and " + std::to_string(kMaxSlots) + " slots");
    // every slot builds its device context on its own thread; creation fails as a whole
    std::vector<std::string> errors(slots);
    uint32_t ready = 0;
    for (uint32_t s = 0; s < slots; ++s)
      workers_.emplace_back([this, s, &errors, &ready] {
        std::unique_ptr<hip::HipBackend> b;
        try {
          b = std::make_unique<hip::HipBackend>(device_);
        } catch (const std::exception& e) {
          errors[s] = e.what();
        }
        {
          std::lock_guard<std::mutex> lk(mu_);
          ++ready;
        }
        ready_cv_.notify_all();
        if (b) run(s, *b);
      });
    {
      std::unique_lock<std::mutex> lk(mu_);
      ready_cv_.wait(lk, [&] { return ready == slots; });
    }
    for (const std::string& e : errors)
      if (!e.empty()) {
        stop();
        throw Error("hip", e);
      }
  }

  ~Server() { stop(); }

  uint64_t submit(const fw_plan* plan, const fw_grid* fixed_grid, const float* const* fixed, const fw_grid* moving_grid,
                  const float* const* moving, double* forward, double* inverse) {
    auto r = std::make_shared<Request>();
    const RegisterPlan p = plan_from(plan);  // validates the plan
    r->plan = *plan;
    r->fixed_grid = grid_of(fixed_grid);
    r->moving_grid = grid_of(moving_grid);
    if (!fixed || !moving) invalid("server", "null image list");
    for (int32_t c = 0; c < plan->channels; ++c) {
      if (!fixed[c] || !moving[c]) invalid("server", "null image data");
      r->fixed.push_back(fixed[c]);
      r->moving.push_back(moving[c]);
    }
    r->forward = forward;
    r->inverse = inverse;
    r->footprint =
        plan_capacity(p, r->fixed_grid, r->moving_grid, size_t(plan->channels)).peak_bytes + AdmissionQueue::kInstanceOverheadBytes;
    std::lock_guard<std::mutex> lk(mu_);
    if (stopping_) invalid("server", "the server is shutting down");
    if (!queue_.fits_alone(r->footprint))
      invalid("server", "the request needs " + std::to_string(r->footprint >> 20) + " MiB of device memory, beyond the budget of " +
                            std::to_string(queue_.budget() >> 20) + " MiB");
    r->id = next_id_++;
    r->submitted = Clock::now();
    requests_[r->id] = r;
    queue_.push(r->id, r->footprint);
    work_cv_.notify_one();
    return r->id;
  }

  // Blocks until the request has ended, then forgets it.
  std::shared_ptr<Request> wait(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) invalid("server", "unknown ticket " + std::to_string(id));
    std::shared_ptr<Request> r = it->second;
    done_cv_.wait(lk, [&] { return r->done; });
    requests_.erase(id);
    return r;
  }

  // A waiting request ends at once as cancelled; a running one stops at its next level or
  // iteration boundary.
  void cancel(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) invalid("server", "unknown ticket " + std::to_string(id));
    Request& r = *it->second;
    if (r.done) return;
    if (queue_.remove(id)) {
      finish_locked(r, FW_ERR_CANCELLED, "cancelled: the request was cancelled before it started");
      return;
    }
    r.cancel = 1;
  }

  void counts(uint32_t* waiting, uint32_t* running, uint64_t* in_use) {
    std::lock_guard<std::mutex> lk(mu_);
    if (waiting) *waiting = uint32_t(queue_.waiting());
    if (running) *running = uint32_t(queue_.running());
    if (in_use) *in_use = queue_.in_use();
  }

 private:
  static constexpr uint32_t kMaxSlots = 16;

  void finish_locked(Request& r, int status, const std::string& error) {
    r.status = status;
    r.error = error;
    r.finished = Clock::now();
    if (r.started == Clock::time_point{}) r.started = r.finished;
    r.done = true;
    done_cv_.notify_all();
  }

  void run(uint32_t slot, hip::HipBackend& b) {
    for (;;) {
      std::shared_ptr<Request> r;
      {
        std::unique_lock<std::mutex> lk(mu_);
        work_cv_.wait(lk, [&] { return stopping_ || queue_.can_admit(); });
        if (stopping_) return;
        r = requests_.at(queue_.admit());
        r->started = Clock::now();
        r->slot = int32_t(slot);
      }
      int status = FW_OK;
      std::string error;
      try {
        execute(b, *r);
      } catch (const Cancelled& e) {
        status = FW_ERR_CANCELLED;
        error = e.what();
      } catch (const Error& e) {
        status = e.what_ == "hip" ? FW_ERR_DEVICE : FW_ERR_INVALID;
        error = e.what();
      } catch (const std::exception& e) {
        status = FW_ERR_INVALID;
        error = e.what();
      }
      // a slot keeps its pool for the next request of the same size, but not beyond the
      // footprint it was admitted with
      if (b.ctx().held() > r->footprint) b.release_idle_buffers();
      std::lock_guard<std::mutex> lk(mu_);
      queue_.finish(r->id);
      finish_locked(*r, status, error);
      work_cv_.notify_all();
    }
  }

  static void execute(hip::HipBackend& b, Request& r) {
    const RegisterPlan p = plan_from(&r.plan);
    std::vector<hip::HImage> fixed, moving;
    for (size_t c = 0; c < r.fixed.size(); ++c) {
      fixed.push_back(b.upload_image_f32(r.fixed_grid, r.fixed[c]));
      moving.push_back(b.upload_image_f32(r.moving_grid, r.moving[c]));
    }
    std::vector<const hip::HImage*> f, m;
    for (size_t c = 0; c < fixed.size(); ++c) {
      f.push_back(&fixed[c]);
      m.push_back(&moving[c]);
    }
    hip::HipLinearEngine engine(b);
    std::optional<hip::HField> fwd, inv;
    const RegisterOutput o = register_resident(b, engine, p, f, m, &r.cancel, r.forward ? &fwd : nullptr,
                                               r.inverse ? &inv : nullptr);
    if (r.forward && fwd) b.download_field_to(*fwd, r.forward);
    if (r.inverse && inv) b.download_field_to(*inv, r.inverse);
    result_to(o, &r.result);
  }

  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stopping_ && workers_.empty()) return;
      stopping_ = true;
      for (auto& kv : requests_) {
        Request& r = *kv.second;
        if (r.done) continue;
        if (queue_.remove(r.id)) finish_locked(r, FW_ERR_CANCELLED, "cancelled: the server shut down");
        else r.cancel = 1;
      }
    }
    work_cv_.notify_all();
    for (std::thread& t : workers_)
      if (t.joinable()) t.join();
    workers_.clear();
  }

  int device_;
  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_, ready_cv_;
  AdmissionQueue queue_;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests_;
  std::vector<std::thread> workers_;
  bool stopping_ = false;
  uint64_t next_id_ = 1;
};

}  // namespace fw::serve

struct fw_server {
  std::unique_ptr<fw::serve::Server> s;
};

namespace {

thread_local std::string g_server_error;

template <typename F>
int server_guard(F&& f) {
  try {
    f();
    g_server_error.clear();
    return FW_OK;
  } catch (const fw::Cancelled& e) {
    g_server_error = e.what();
    return FW_ERR_CANCELLED;
  } catch (const fw::Error& e) {
    g_server_error = e.what();
    return e.what_ == "hip" ? FW_ERR_DEVICE : FW_ERR_INVALID;
  } catch (const std::exception& e) {
    g_server_error = e.what();
    return FW_ERR_INVALID;
  }
}

}  // namespace

extern "C" {

int fw_server_create(int32_t device, uint32_t slots, uint64_t budget_bytes, fw_server** out) {
  return server_guard([&] {
    if (!out) fw::invalid("server", "null output");
    *out = nullptr;
    uint64_t budget = budget_bytes;
    if (budget == 0) budget = fw::hip::HipBackend(device).memory_budget();
    *out = new fw_server{std::make_unique<fw::serve::Server>(device, slots == 0 ? 4u : slots, budget)};
  });
}

void fw_server_destroy(fw_server* server) { delete server; }

const char* fw_server_last_error(void) { return g_server_error.c_str(); }

int fw_server_submit(fw_server* server, const fw_plan* plan, const fw_grid* fixed_grid, const float* const* fixed,
                     const fw_grid* moving_grid, const float* const* moving, double* forward, double* inverse,
                     uint64_t* ticket) {
  return server_guard([&] {
    if (!server || !ticket) fw::invalid("server", "null server or ticket");
    *ticket = server->s->submit(plan, fixed_grid, fixed, moving_grid, moving, forward, inverse);
  });
}

int fw_server_wait(fw_server* server, uint64_t ticket, fw_result* result, fw_served_timing* timing) {
  std::shared_ptr<fw::serve::Request> r;
  const int st = server_guard([&] {
    if (!server) fw::invalid("server", "null server");
    r = server->s->wait(ticket);
  });
  if (st != FW_OK) return st;
  if (timing) {
    auto secs = [](fw::serve::Clock::time_point a, fw::serve::Clock::time_point b) {
      return std::chrono::duration<double>(b - a).count();
    };
    timing->queued_s = secs(r->submitted, r->started);
    timing->run_s = secs(r->started, r->finished);
    timing->total_s = secs(r->submitted, r->finished);
    timing->slot = r->slot;
    timing->footprint_bytes = r->footprint;
  }
  if (result) *result = r->result;
  g_server_error = r->error;
  return r->status;
}

int fw_server_cancel(fw_server* server, uint64_t ticket) {
  return server_guard([&] {
    if (!server) fw::invalid("server", "null server");
    server->s->cancel(ticket);
  });
}

int fw_server_counts(fw_server* server, uint32_t* waiting, uint32_t* running, uint64_t* bytes_in_use) {
  return server_guard([&] {
    if (!server) fw::invalid("server", "null server");
    server->s->counts(waiting, running, bytes_in_use);
  });
}

}  // extern "C"
