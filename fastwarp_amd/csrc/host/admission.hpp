// Admission of registration requests to a server's slots by device-memory footprint.
//
// The reference runs several registrations on one GPU as independent device instances and sizes
// their number with its capacity model (crates/fastwarp-cli/src/engine.rs:1015-1068
// `registrations_at_once`: each instance's plan must fit an equal share of the budget after a
// per-instance overhead, engine.rs:1000 DEVICE_OVERHEAD_BYTES). A server admits requests one at a
// time instead, as slots and memory free up: a request's footprint is its plan's peak bytes plus
// that overhead, and the head of the queue starts when an idle slot asks for work and the
// footprint fits beside what is held already. Requests start in arrival order; a head that does
// not fit yet holds the ones behind it back, so a large request is never starved.
//
// What is held: the footprints of the running requests, and the idle buffers a slot's pool keeps
// between requests (`retained`). A slot that starts a request still holds the buffers it kept, and
// its new request may not reuse them (the pool reuses buffers by size), so a running request is
// charged its footprint plus what its slot kept when it started. A slot that keeps buffers while
// the head does not fit gives them back (the request loop does that), so a head is never held
// back by idle memory alone.
#pragma once

#include <cstdint>
#include <deque>
#include <unordered_map>

namespace fw {

class AdmissionQueue {
 public:
  // engine.rs:1000: device bytes of one instance beyond its plan (context, pools, staging)
  static constexpr uint64_t kInstanceOverheadBytes = uint64_t(256) << 20;

  explicit AdmissionQueue(uint64_t budget) : budget_(budget) {}

  uint64_t budget() const { return budget_; }
  // A request whose footprint exceeds the whole budget can never start.
  bool fits_alone(uint64_t footprint) const { return footprint <= budget_; }

  void push(uint64_t id, uint64_t footprint) { waiting_.push_back({id, footprint}); }

  bool has_waiting() const { return !waiting_.empty(); }
  // Whether the head of the queue may start now beside everything held.
  bool can_admit() const { return !waiting_.empty() && waiting_.front().footprint <= budget_ - held(); }

  // Starts the head (can_admit() must hold) on a slot that keeps `retained` idle bytes (counted
  // in with retain() before) and returns its id.
  uint64_t admit(uint64_t retained = 0) {
    const Entry e = waiting_.front();
    waiting_.pop_front();
    const uint64_t r = retained <= idle_ ? retained : idle_;
    idle_ -= r;
    in_use_ += e.footprint + r;
    running_[e.id] = e.footprint + r;
    return e.id;
  }

  // A running request ended: its charge is free again.
  void finish(uint64_t id) {
    auto it = running_.find(id);
    if (it == running_.end()) return;
    in_use_ -= it->second;
    running_.erase(it);
  }

  // An idle slot's pool now keeps `now` bytes instead of `before` (both as that slot reported
  // them: after a request ends `before` is 0, after it releases its buffers `now` is).
  void retain(uint64_t before, uint64_t now) {
    idle_ -= before <= idle_ ? before : idle_;
    idle_ += now;
  }

  // Drops a request that has not started; false when it is not waiting.
  bool remove(uint64_t id) {
    for (auto it = waiting_.begin(); it != waiting_.end(); ++it)
      if (it->id == id) {
        waiting_.erase(it);
        return true;
      }
    return false;
  }

  size_t waiting() const { return waiting_.size(); }
  size_t running() const { return running_.size(); }
  // Bytes charged to running requests, and bytes idle slots keep.
  uint64_t in_use() const { return in_use_; }
  uint64_t idle() const { return idle_; }
  uint64_t held() const { return in_use_ + idle_ <= budget_ ? in_use_ + idle_ : budget_; }

 private:
  struct Entry {
    uint64_t id, footprint;
  };
  uint64_t budget_;
  uint64_t in_use_ = 0, idle_ = 0;
  std::deque<Entry> waiting_;
  std::unordered_map<uint64_t, uint64_t> running_;
};

}  // namespace fw
