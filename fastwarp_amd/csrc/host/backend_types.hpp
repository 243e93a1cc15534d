// Value types every backend of the SyN loop shares: the inversion parameters and statistics
// (crates/fastwarp-reg/src/field_ops.rs:276-307) and a metric evaluated in both directions.
#pragma once

#include <cstdint>

namespace fw {

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
};

}  // namespace fw
