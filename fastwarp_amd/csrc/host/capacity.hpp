// Pre-flight device-memory model of one registration on this engine: the role of the reference's
// capacity model (crates/fastwarp-gpu/src/capacity.rs:162-180 `memory_budget`, :511-609
// `plan_capacity_syn_masked`, :715- `plan_capacity_linear`) for this engine's own buffers. Pure
// and device-free: from the plan and the two grids it gives the bytes the linear chain and the SyN
// stage hold at their peak, so a caller (an admission queue) can decide before any image is
// loaded whether a registration fits a memory budget, and how many fit at once.
//
// The per-voxel figures are this engine's buffer-pool high-water marks measured on the 2 mm and
// 1 mm cohort workloads (1 and 2 channels; identical per voxel at both sizes), rounded up:
//   smoothing cache  4 B per voxel per image per distinct sigma (sigma 0 is the upload itself);
//   linear chain     per sample of the finest level 40 B (16 B record, 16 B continuous index,
//                    8 B of values) plus 15 % for the coarser levels' records kept by the chain;
//   SyN stage        105 B per voxel of the finest level for the first channel (four total
//                    fields, the warped pair, both derivative fields, the smoothing output, the
//                    composition and inversion temporaries and the idle buffers of coarser
//                    levels), 14 B for every further channel.
// Each figure gets 5 % of headroom, and every registration 16 MiB for scratch, taps and events.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <set>

#include "register.hpp"

namespace fw {

struct CapacityPlan {
  uint64_t fixed_voxels = 0, moving_voxels = 0;
  uint64_t smoothing_cache_bytes = 0;  // both images of every channel at every distinct sigma
  uint64_t linear_bytes = 0;           // peak of the linear chain, its cache included (0 if not run)
  uint64_t syn_bytes = 0;              // peak of the SyN stage, the whole cache included (0 if not run)
  uint64_t peak_bytes = 0;             // the larger of the two, plus the fixed overhead
  uint64_t largest_buffer_bytes = 0;   // the largest single allocation
};

constexpr uint64_t kCapacityOverheadBytes = uint64_t(16) << 20;
constexpr double kCapacityHeadroom = 1.05;
constexpr double kSynBytesFirstChannel = 105.0;
constexpr double kSynBytesPerExtraChannel = 14.0;
constexpr double kLinearBytesPerSample = 40.0;
constexpr double kLinearCoarserLevels = 1.15;

inline uint64_t capacity_round(double bytes) { return uint64_t(std::ceil(bytes * kCapacityHeadroom)); }

inline CapacityPlan plan_capacity(const RegisterPlan& plan, const Grid& fixed, const Grid& moving, size_t channels) {
  if (channels == 0) invalid("capacity", "channels must be at least 1");
  CapacityPlan c;
  c.fixed_voxels = fixed.num_voxels();
  c.moving_voxels = moving.num_voxels();
  const double nf = double(c.fixed_voxels), nm = double(c.moving_voxels);
  // distinct sigmas per channel: the first channel serves both stages, the others only SyN
  std::set<double> first{0.0}, rest{0.0};
  if (plan.run_linear)
    for (double s : plan.rigid.smoothing_sigmas) first.insert(std::max(s, 0.0));
  if (plan.run_linear)
    for (double s : plan.affine.smoothing_sigmas) first.insert(std::max(s, 0.0));
  if (plan.run_syn)
    for (double s : plan.syn.smoothing_sigmas) {
      first.insert(std::max(s, 0.0));
      rest.insert(std::max(s, 0.0));
    }
  std::set<double> linear_sigmas{0.0};
  for (double s : plan.rigid.smoothing_sigmas) linear_sigmas.insert(std::max(s, 0.0));
  for (double s : plan.affine.smoothing_sigmas) linear_sigmas.insert(std::max(s, 0.0));
  const double pair = 4.0 * (nf + nm);
  const double cache = pair * (double(first.size()) + double(channels - 1) * double(rest.size()));
  c.smoothing_cache_bytes = uint64_t(cache);
  if (plan.run_linear) {
    const double samples = nf * std::min(1.0, double(plan.rigid.sampling_percentage));
    c.linear_bytes = capacity_round(pair * double(linear_sigmas.size()) + samples * kLinearBytesPerSample * kLinearCoarserLevels);
  }
  if (plan.run_syn) {
    const double working = nf * (kSynBytesFirstChannel + kSynBytesPerExtraChannel * double(channels - 1));
    c.syn_bytes = capacity_round(cache + working);
  }
  c.peak_bytes = std::max(c.linear_bytes, c.syn_bytes) + kCapacityOverheadBytes;
  c.largest_buffer_bytes = uint64_t(std::max(12.0 * nf, 4.0 * nm));
  return c;
}

// How many registrations of this plan fit `budget` bytes at once (at most `max`, at least 0).
inline uint32_t registrations_at_once(const CapacityPlan& c, uint64_t budget, uint32_t max) {
  if (c.peak_bytes == 0) return max;
  return uint32_t(std::min<uint64_t>(max, budget / c.peak_bytes));
}

}  // namespace fw
