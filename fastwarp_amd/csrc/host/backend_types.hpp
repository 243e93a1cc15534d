// Value types every backend of the SyN loop shares: the inversion parameters and statistics
// (crates/fastwarp-reg/src/field_ops.rs:276-307) and a metric evaluated in both directions.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <utility>

namespace fw {

// A result a backend reads back from the device later: get() is valid once the backend has
// settled (settle() waits for the device and for every read-back issued since the last settle).
// Backends that compute on the host hand out results that are ready at once.
template <typename T>
class Pending {
 public:
  Pending() = default;
  explicit Pending(T value) : value_(std::move(value)), ready_(true) {}
  explicit Pending(std::function<T()> read) : read_(std::move(read)) {}
  T get() {
    if (!ready_) {
      value_ = read_();
      ready_ = true;
      read_ = nullptr;
    }
    return value_;
  }

 private:
  T value_{};
  bool ready_ = false;
  std::function<T()> read_;
};

// InvertParams::SYN by default (field_ops.rs:290-295).
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

// (fixed, moving) then (moving, fixed): values, derivative fields and, when asked, their l2 norms.
template <typename F>
struct MetricPair {
  double values[2] = {0, 0};
  F fields[2];
  double norms[2] = {0, 0};
  // Set instead of `values` by the deferred evaluation: both values, once the backend settled.
  Pending<std::array<double, 2>> later;
};

}  // namespace fw
