// The registration server of include/fastwarp_hip.h: continuous batching of whole registrations
// on one device. Each slot is a worker thread with its own HipBackend (its own HIP stream and
// buffer pool), the reference's model of several registrations on one GPU as independent device
// instances (crates/fastwarp-cli/src/engine.rs:1015-1068). Requests are admitted by device-memory
// footprint as slots and memory free up (host/admission.hpp), their kernels overlap on the device
// through the slots' streams, and a request's result does not depend on what runs beside it: a
// slot's work reads and writes only its own buffers, and every reduction has a fixed order.
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../../include/fastwarp_hip.h"
#include "hip/backend.hpp"
#include "host/admission.hpp"
#include "host/capacity.hpp"
#include "host/request_loop.hpp"
#include "plan.hpp"

namespace fw::serve {

// One registration request: the caller's arrays (valid until it is waited for) and its result.
struct Job {
  fw_plan plan{};
  Grid fixed_grid, moving_grid;
  std::vector<const float*> fixed, moving;
  double* forward = nullptr;
  double* inverse = nullptr;
  fw_result result{};
};

// Uploads the request's images on the slot's stream, registers them on the device and downloads
// the fields the caller asked for.
inline void execute(hip::HipBackend& b, Job& r, const volatile int32_t* cancel) {
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
  const RegisterOutput o = register_resident(b, engine, p, f, m, cancel, r.forward ? &fwd : nullptr, r.inverse ? &inv : nullptr);
  if (r.forward && fwd) b.download_field_to(*fwd, r.forward);
  if (r.inverse && inv) b.download_field_to(*inv, r.inverse);
  result_to(o, &r.result);
}

class Server {
 public:
  Server(int device, uint32_t slots, uint64_t budget)
      : loop_(
            check_slots(slots), budget, [device](uint32_t) { return std::make_unique<hip::HipBackend>(device); },
            [](hip::HipBackend& b, Job& job, const volatile int32_t* cancel) {
              // the slot keeps its pool for the next request of the same size, but not beyond the
              // footprint this one was admitted with, and not at all after a failure
              try {
                execute(b, job, cancel);
              } catch (...) {
                b.release_idle_buffers();
                throw;
              }
              if (b.ctx().held() > footprint_of(job)) b.release_idle_buffers();
            },
            [](hip::HipBackend& b) { return uint64_t(b.ctx().held()); },
            [](hip::HipBackend& b) { b.release_idle_buffers(); }) {}

  uint64_t submit(const fw_plan* plan, const fw_grid* fixed_grid, const float* const* fixed, const fw_grid* moving_grid,
                  const float* const* moving, double* forward, double* inverse) {
    Job j;
    plan_from(plan);  // validates the plan
    j.plan = *plan;
    j.fixed_grid = grid_of(fixed_grid);
    j.moving_grid = grid_of(moving_grid);
    if (!fixed || !moving) invalid("server", "null image list");
    for (int32_t c = 0; c < plan->channels; ++c) {
      if (!fixed[c] || !moving[c]) invalid("server", "null image data");
      j.fixed.push_back(fixed[c]);
      j.moving.push_back(moving[c]);
    }
    j.forward = forward;
    j.inverse = inverse;
    const uint64_t footprint = footprint_of(j);
    return loop_.submit(footprint, std::move(j));
  }

  std::pair<Outcome, Job> wait(uint64_t id) { return loop_.wait(id); }
  void cancel(uint64_t id) { loop_.cancel(id); }
  void counts(uint32_t* waiting, uint32_t* running, uint64_t* in_use) { loop_.counts(waiting, running, in_use); }

 private:
  static constexpr uint32_t kMaxSlots = 16;
  static uint32_t check_slots(uint32_t slots) {
    if (slots < 1 || slots > kMaxSlots) invalid("server", "between 1 and " + std::to_string(kMaxSlots) + " slots");
    return slots;
  }
  // the capacity model's peak of the plan plus the instance overhead
  static uint64_t footprint_of(const Job& j) {
    return plan_capacity(plan_from(&j.plan), j.fixed_grid, j.moving_grid, size_t(j.plan.channels)).peak_bytes +
           AdmissionQueue::kInstanceOverheadBytes;
  }
  RequestLoop<hip::HipBackend, Job> loop_;
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
  std::pair<fw::Outcome, fw::serve::Job> r;
  const int st = server_guard([&] {
    if (!server) fw::invalid("server", "null server");
    r = server->s->wait(ticket);
  });
  if (st != FW_OK) return st;
  const fw::Outcome& o = r.first;
  if (timing) {
    timing->queued_s = o.queued_s;
    timing->run_s = o.run_s;
    timing->total_s = o.total_s;
    timing->slot = o.slot;
    timing->footprint_bytes = o.footprint;
  }
  if (result) *result = r.second.result;
  g_server_error = o.error;
  switch (o.status) {
    case fw::Ended::Ok: return FW_OK;
    case fw::Ended::Cancelled: return FW_ERR_CANCELLED;
    case fw::Ended::Device: return FW_ERR_DEVICE;
    default: return FW_ERR_INVALID;
  }
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
