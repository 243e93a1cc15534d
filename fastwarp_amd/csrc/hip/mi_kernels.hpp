// Launch interface of the Mattes mutual-information kernels (linear-stage sampled MI and the dense
// SyN-stage pair). Histograms are exact integers: Parzen weights in 2^-32 fixed point summed with
// 64-bit LDS atomics, so the result does not depend on the order of the additions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fw::hip {

constexpr int kMiMaxBins = 64;
constexpr double kMiFixedPoint = 4294967296.0;  // 2^32

// One linear-stage evaluation (crates/fastwarp-reg/src/metric/mi.rs:364-437 `accumulate`).
struct MiSampledArgs {
  const float4* samples;  // (x, y, z) = physical point minus `anchor`, w = fixed bin (as int bits)
  int count;
  const float* moving;
  Dims md;
  float a[9];             // moving continuous index = a * s + b
  float b[3];
  float inv_bin_size;     // 1 / moving_bin_size
  float moving_min, moving_max;
  int bins;
  unsigned long long* hist;  // bins*bins joint, then bins fixed marginal (counts), then the valid count
  unsigned long long* hist_slots;  // per block of the histogram pass: one such histogram
  // derivative pass only
  const float* ratio;     // bins*bins weights
  Probes probes;          // central-difference probes of the moving grid
  float outer[9];         // outer matrix that left-multiplies the Jacobian columns
  float rel_center[3];    // anchor - optimised centre (physical)
  float trig[6];          // cx sx cy sy cz sz of the Euler angles
  int kind;               // 0 Euler3D (Rz Rx Ry), 1 affine
  double* partial;        // per block: 12 doubles
  int threads;            // threads per block of the histogram passes (0: mi_sampled_threads)
};
void launch_mi_sampled_hist(const MiSampledArgs& a, int blocks, hipStream_t s);
// Several transforms over the same samples in one pass (a.a / a.b unused): transform k maps the
// samples with batch.a[k], batch.b[k]; block b's slot holds one histogram per transform, in order.
constexpr int kMiMaxBatch = 4;
struct MiBatch {
  float a[kMiMaxBatch][9];
  float b[kMiMaxBatch][3];
};
// The number of transforms one batched pass takes for `bins` (its LDS holds a histogram each).
int mi_batch_capacity(int bins);
void launch_mi_sampled_hist_batch(const MiSampledArgs& a, const MiBatch& batch, int count, int blocks, hipStream_t s);
// Adds `nslots` per-block histograms of `cells` u64 entries into `hist`.
// hist[c] = the slots' cell c added up; block b then stores `tag` into done[b] with a system-scope
// release, b < mi_sum_slots_blocks(cells), unless done is null (sums left for a later kernel).
int mi_sum_slots_blocks(int cells);
void launch_mi_sum_slots(const unsigned long long* slots, int nslots, int cells, unsigned long long* hist, uint32_t* done,
                         uint32_t tag, hipStream_t s);
// The Mattes finish of `count` histograms laid out as k_mi_sum_slots writes them (each bins*bins
// joint cells, bins fixed marginal counts, the valid count), on the device: out[k] receives the
// value of histogram k, its valid count and a status (0 ok, 1 no valid sample, 2 the joint cells
// sum to zero); `ratio` (if not null) the derivative pass's f32 weights of histogram 0; then, if
// done is not null, done[k] = tag (system-scope release), k < count.
struct MiFinishOut {
  double value;
  unsigned long long valid;
  int32_t status;
  int32_t pad;
};
void launch_mi_finish(const unsigned long long* hist, int count, int bins, double moving_bin_size, MiFinishOut* out,
                      float* ratio, uint32_t* done, uint32_t tag, hipStream_t s);
// launch_mi_sum_slots and launch_mi_finish in one launch, for `count` histograms of at most
// kMiFusedBins bins in `nslots` slots (slot q: the count histograms one after another): each block
// adds its histogram's slots in LDS and finishes it. For passes of few slots, where the two
// launches cost more than the sums.
constexpr int kMiFusedBins = 32;
constexpr int kMiFusedSlots = 32;
void launch_mi_sum_finish(const unsigned long long* slots, int nslots, int count, int bins, double moving_bin_size,
                          MiFinishOut* out, float* ratio, uint32_t* done, uint32_t tag, hipStream_t s);
void launch_mi_sampled_deriv(const MiSampledArgs& a, int blocks, hipStream_t s);
int mi_sampled_blocks(int count, int cap = 512);
int mi_sampled_threads(int count);
constexpr int kMiDerivMaxBlocks = 2048;
int mi_deriv_blocks(int count, int cap = kMiDerivMaxBlocks);
// out[k] (k < np) = the blocks' derivative partials added in a fixed order; *finish_out = *finish
// if finish is given (a finish left in device memory, passed to the host with the sums); then
// done[0] = tag (system-scope release).
void launch_mi_deriv_finish(const double* partial, int blocks, int np, double* out, uint32_t* done, uint32_t tag,
                            hipStream_t s, const MiFinishOut* finish = nullptr, MiFinishOut* finish_out = nullptr);

// Dense pair on one grid (mi.rs:578-676 in both directions, wgpu_backend/mi/dense.rs:241-487).
struct MiDenseArgs {
  const float* f;
  const float* m;
  const float* fmask;  // may be null
  const float* mmask;  // may be null
  Dims d;
  int bins;
  // direction 0: fixed = f, moving = m; direction 1: fixed = m, moving = f
  float fixed_thr[2][kMiMaxBins];  // f32 thresholds of the fixed-axis bins 3..bins-3
  float moving_min[2], moving_max[2], inv_bin_size[2];
  unsigned long long* hist;        // 2 x (bins*bins + bins + 1)
  const float* ratio;              // 2 x bins*bins
  Probes probes;
  float* d0;
  float* d1;
  float* range_partial;            // per block: fmin fmax mmin mmax
};
void launch_mi_dense_range(const MiDenseArgs& a, int blocks, hipStream_t s);
void launch_mi_dense_hist(const MiDenseArgs& a, int blocks, hipStream_t s);
void launch_mi_dense_deriv(const MiDenseArgs& a, hipStream_t s);
int mi_dense_blocks(long long n);

}  // namespace fw::hip
