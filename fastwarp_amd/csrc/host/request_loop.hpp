// The request loop of the registration server: slot threads that take queued requests in arrival
// order as slots and device memory free up (admission.hpp), run them, and hand each result to the
// one caller that waits for it. The reference runs several registrations on one GPU as independent
// device instances on worker threads (python/fastwarp/_api.py:1104-1130, sized by
// crates/fastwarp-cli/src/engine.rs:1015-1068 `registrations_at_once`); here the instances are the
// slots' own `Slot` objects, each built on its own thread.
//
// Slot: whatever one thread needs to run requests (a device context); built by `make_slot(i)` on
// slot thread i. Payload: a request's inputs and outputs, owned by the loop from submission until
// it has been waited for. `run(slot, payload, cancel)` runs one request; it may throw
// (Cancelled, Error or any std::exception) and should poll `*cancel` with an atomic load (the
// loop sets it from another thread with an atomic store). `held(slot)` is the device bytes the
// slot's pool keeps once a request has ended, and `release(slot)` gives them back; the loop
// charges what a slot keeps against the budget (admission.hpp) and has an idle slot release it
// whenever the head of the queue does not fit.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "admission.hpp"
#include "common.hpp"

namespace fw {

enum class Ended { Ok, Invalid, Device, Cancelled };

struct Outcome {
  Ended status = Ended::Ok;
  std::string error;
  double queued_s = 0.0, run_s = 0.0, total_s = 0.0;
  int32_t slot = -1;  // -1: it never started
  uint64_t footprint = 0;
};

template <typename Slot, typename Payload>
class RequestLoop {
 public:
  using MakeSlot = std::function<std::unique_ptr<Slot>(uint32_t)>;
  using Run = std::function<void(Slot&, Payload&, const volatile int32_t*)>;
  using Held = std::function<uint64_t(Slot&)>;
  using Release = std::function<void(Slot&)>;

  // Starts `slots` threads; throws (Error "hip" with the first message) when a slot cannot be
  // built, after stopping the others.
  RequestLoop(uint32_t slots, uint64_t budget, MakeSlot make_slot, Run run, Held held, Release release)
      : queue_(budget), run_(std::move(run)), held_(std::move(held)), release_(std::move(release)) {
    std::vector<std::string> errors(slots);
    uint32_t ready = 0;
    for (uint32_t s = 0; s < slots; ++s)
      threads_.emplace_back([this, s, &errors, &ready, &make_slot] {
        std::unique_ptr<Slot> slot;
        try {
          slot = make_slot(s);
        } catch (const std::exception& e) {
          errors[s] = e.what();
        }
        {
          std::lock_guard<std::mutex> lk(mu_);
          ++ready;
        }
        ready_cv_.notify_all();
        if (slot) serve(s, *slot);
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

  ~RequestLoop() { stop(); }
  RequestLoop(const RequestLoop&) = delete;
  RequestLoop& operator=(const RequestLoop&) = delete;

  // Queues a request; throws Error when its footprint exceeds the whole budget or the loop stops.
  uint64_t submit(uint64_t footprint, Payload payload) {
    auto r = std::make_shared<Request>();
    r->footprint = footprint;
    r->payload = std::move(payload);
    std::lock_guard<std::mutex> lk(mu_);
    if (stopping_) invalid("server", "the server is shutting down");
    if (!queue_.fits_alone(footprint))
      invalid("server", "the request needs " + std::to_string(footprint >> 20) + " MiB of device memory, beyond the budget of " +
                            std::to_string(queue_.budget() >> 20) + " MiB");
    r->id = next_id_++;
    r->submitted = Clock::now();
    requests_[r->id] = r;
    queue_.push(r->id, footprint);
    work_cv_.notify_one();
    return r->id;
  }

  // Blocks until the request has ended and hands it back (once: the ticket is unknown to every
  // later or concurrent wait).
  std::pair<Outcome, Payload> wait(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end() || it->second->claimed) invalid("server", "unknown ticket " + std::to_string(id));
    std::shared_ptr<Request> r = it->second;
    r->claimed = true;
    done_cv_.wait(lk, [&] { return r->done; });
    requests_.erase(id);
    Outcome o;
    o.status = r->status;
    o.error = r->error;
    o.queued_s = seconds(r->submitted, r->started);
    o.run_s = seconds(r->started, r->finished);
    o.total_s = seconds(r->submitted, r->finished);
    o.slot = r->slot;
    o.footprint = r->footprint;
    return {std::move(o), std::move(r->payload)};
  }

  // A waiting request ends at once; a running one is asked to stop (its `cancel` word is set).
  void cancel(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) invalid("server", "unknown ticket " + std::to_string(id));
    Request& r = *it->second;
    if (r.done) return;
    if (queue_.remove(id)) {
      finish_locked(r, Ended::Cancelled, "cancelled: the request was cancelled before it started");
      return;
    }
    __atomic_store_n(&r.cancel, 1, __ATOMIC_RELAXED);
  }

  void counts(uint32_t* waiting, uint32_t* running, uint64_t* in_use) {
    std::lock_guard<std::mutex> lk(mu_);
    if (waiting) *waiting = uint32_t(queue_.waiting());
    if (running) *running = uint32_t(queue_.running());
    if (in_use) *in_use = queue_.in_use();
  }

 private:
  using Clock = std::chrono::steady_clock;
  struct Request {
    uint64_t id = 0, footprint = 0;
    Payload payload{};
    volatile int32_t cancel = 0;
    Clock::time_point submitted, started, finished;
    int32_t slot = -1;
    bool done = false;
    bool claimed = false;  // a caller waits for it
    Ended status = Ended::Ok;
    std::string error;
  };

  static double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

  void finish_locked(Request& r, Ended status, const std::string& error) {
    r.status = status;
    r.error = error;
    r.finished = Clock::now();
    if (r.started == Clock::time_point{}) r.started = r.finished;
    r.done = true;
    done_cv_.notify_all();
  }

  void serve(uint32_t index, Slot& slot) {
    uint64_t kept = 0;          // idle bytes this slot's pool keeps, as charged to the queue
    bool released = false;      // nothing more to give back until the next request
    for (;;) {
      std::shared_ptr<Request> r;
      {
        std::unique_lock<std::mutex> lk(mu_);
        work_cv_.wait(lk, [&] {
          return stopping_ || queue_.can_admit() || (kept > 0 && !released && queue_.has_waiting());
        });
        if (stopping_) {
          queue_.retain(kept, 0);
          return;
        }
        if (!queue_.can_admit()) {
          // the head does not fit beside what is held: give this slot's idle buffers back
          lk.unlock();
          uint64_t now = kept;
          try {
            release_(slot);
            now = held_(slot);
          } catch (const std::exception&) {
          }
          lk.lock();
          queue_.retain(kept, now);
          kept = now;
          released = true;
          work_cv_.notify_all();
          continue;
        }
        r = requests_.at(queue_.admit(kept));
        kept = 0;
        released = false;
        r->started = Clock::now();
        r->slot = int32_t(index);
      }
      Ended status = Ended::Ok;
      std::string error;
      try {
        run_(slot, r->payload, &r->cancel);
      } catch (const Cancelled& e) {
        status = Ended::Cancelled;
        error = e.what();
      } catch (const Error& e) {
        status = e.what_ == "hip" ? Ended::Device : Ended::Invalid;
        error = e.what();
      } catch (const std::exception& e) {
        status = Ended::Invalid;
        error = e.what();
      }
      uint64_t now = 0;
      try {
        now = held_(slot);
      } catch (const std::exception&) {
      }
      std::lock_guard<std::mutex> lk(mu_);
      queue_.finish(r->id);
      queue_.retain(0, now);
      kept = now;
      finish_locked(*r, status, error);
      work_cv_.notify_all();
    }
  }

  // Cancels what has not ended (waiting requests at once, running ones through their words) and
  // joins the slots; requests no one waited for are dropped with the loop.
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stopping_ = true;
      for (auto& kv : requests_) {
        Request& r = *kv.second;
        if (r.done) continue;
        if (queue_.remove(r.id)) finish_locked(r, Ended::Cancelled, "cancelled: the server shut down");
        else __atomic_store_n(&r.cancel, 1, __ATOMIC_RELAXED);
      }
    }
    work_cv_.notify_all();
    for (std::thread& t : threads_)
      if (t.joinable()) t.join();
    threads_.clear();
  }

  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_, ready_cv_;
  AdmissionQueue queue_;
  Run run_;
  Held held_;
  Release release_;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests_;
  std::vector<std::thread> threads_;
  bool stopping_ = false;
  uint64_t next_id_ = 1;
};

}  // namespace fw
