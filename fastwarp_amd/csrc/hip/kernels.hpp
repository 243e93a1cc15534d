// Launch interface of the gfx950 kernels. Host code (the backend) fills these plain structs in f64
// and narrows to f32; the kernels see only raw device pointers and sizes.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

namespace fw::hip {

// Events carried by the launches of an open timing scope on this thread: the scope's first launch
// records `start` when it begins, every launch records `stop` when it ends, so the span is the
// kernels' own execution (hipExtLaunchKernelGGL), with no marker packets between them.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  int launches = 0;
};

// The timing scope open on this thread, or nullptr.
LaunchEvents*& current_launch_events();

// Every kernel of the library is launched through here.
template <typename... Params, typename... Args>
inline void launch(void (*kernel)(Params...), dim3 grid, dim3 block, uint32_t lds, hipStream_t s, Args... args) {
  LaunchEvents* ev = current_launch_events();
  if (ev == nullptr) {
    kernel<<<grid, block, lds, s>>>(args...);
    return;
  }
  hipExtLaunchKernelGGL(kernel, grid, block, lds, s, ev->launches == 0 ? ev->start : nullptr, ev->stop, 0u, args...);
  ++ev->launches;
}

// A continuous-index map c = m * (i, j, k) + c0 (+ l * u for a displacement u), f32 row-major.
struct IndexMap {
  float m[9];
  float c0[3];
  float l[9];
};

struct Dims {
  int nx, ny, nz;
  __host__ __device__ long long n() const { return (long long)nx * ny * nz; }
};

// Offsets (continuous index units) of the half-spacing probes of a central difference along each
// physical axis, and the spacing that divides their difference.
struct Probes {
  float o[3][3];  // o[d] = index offset of +h_d e_d
  float inv_spacing[3];
  int diagonal;   // every offset lies on its own grid axis with magnitude 0.5
};

// --- CC pair (fused window sums, both directions) ---
struct CcArgs {
  const float* fixed;
  const float* moving;
  const float* valid;  // nullptr: every voxel valid
  Dims d;
  int radius;
  int tile_z;
  float scale_f, scale_m;  // 2^19 / (the image's bound rounded up to a power of two)
  float step_f, step_m;    // 1 / scale
  const float* bounds;     // device [2] (fixed, moving): when set, the scales come from these
  Probes probes;
  float* d_fixed;        // 3 planes: coefficient * grad(moving)
  float* d_moving;       // 3 planes: coefficient' * grad(fixed)
  float* partial_cc;     // per block: sum of localCC over its valid voxels
  int* partial_count;    // per block: valid voxels
};
void launch_cc_pair(const CcArgs& a, hipStream_t s);
dim3 cc_grid(const Dims& d, int tile_z);
// The longest z run of at most max_tile planes (halving) whose grid still has min_blocks blocks.
int cc_tile_for(const Dims& d, int max_tile, int min_blocks = 512);
constexpr int kCcTileZ = 16;  // the default longest z run

// --- per-block partial reductions ---
// Largest |u/spacing| of two fields (b may be null) and sums of squares.
void launch_field_norms(const float* a, const float* b, long long n, float3 inv_spacing, float* partial_max,
                        float* partial_sq, int blocks, hipStream_t s);
int norm_blocks(long long n);
// Per-block finite minimum / maximum (+inf / -inf for a block without finite values).
void launch_finite_range(const float* x, long long n, float* partial_min, float* partial_max, int blocks, hipStream_t s);
// Nearest-voxel and linear values at m continuous indices (x, y, z; w unused); NaN outside.
// Device preprocessing (preprocess.rs, init.rs:47-86).
constexpr int kIntensityBins = 256;
void launch_range_and_bad(const float* x, long long n, float* partial_min, float* partial_max, unsigned int* partial_bad,
                          int blocks, hipStream_t s);
void launch_intensity_hist(const float* x, long long n, const float* thresholds, unsigned long long* slots, int blocks,
                           hipStream_t s);
void launch_window(const float* x, float* out, long long n, float lo, float hi, float sh, float sl, float th, float tl,
                   float out_min, float out_max, hipStream_t s);
dim3 moments_grid(const Dims& d);
void launch_moments(const float* x, Dims d, float* partial, hipStream_t s);
void launch_sample_values(const float* img, Dims d, const float4* idx, long long m, float* nearest, float* linear,
                          hipStream_t s);

// --- elementwise ---
void launch_scale(float* x, long long count, float s, hipStream_t st);
// x *= *factor, the factor read on the device.
void launch_scale_by(float* x, long long count, const float* factor, hipStream_t st);
// x *= factors[0], y *= factors[1] (both of `count` floats) in one launch.
void launch_scale_by2(float* x, float* y, long long count, const float* factors, hipStream_t st);
// Per field f < fields: maxima[f] = max of partial_max[f * blocks ...] (NaN propagates),
// factors[f] = step / maxima[f] (step when not positive). One block.
void launch_step_factors(const float* partial_max, int blocks, int fields, double step, float* factors, float* maxima,
                         hipStream_t s);
void launch_axpy(float* acc, const float* x, long long count, float a, hipStream_t st);

// --- warp / resample ---
struct WarpArgs {
  const float* src;
  Dims sd;               // source dims
  const float* field;    // 3 planes on the output grid
  Dims od;               // output dims
  IndexMap to_src;       // output index (+ displacement) -> source continuous index
  const float* pre;      // optional prior field on its own grid
  Dims pd;
  IndexMap to_pre;       // output index (+ displacement) -> prior continuous index
  float pre_l[9];        // prior displacement -> source continuous index
  int nearest;           // 1: mask mode (nearest voxel, {0,1})
  const float* product;  // mask mode: multiply into
  float* out;
};
void launch_warp(const WarpArgs& a, hipStream_t s);
// Two warps with the same output dims in one launch.
void launch_warp2(const WarpArgs& a0, const WarpArgs& a1, hipStream_t s);

struct ResampleArgs {
  const float* src;  // 3 planes
  Dims sd;
  Dims od;
  IndexMap to_src;
  float* out;        // 3 planes
};
void launch_resample(const ResampleArgs& a, hipStream_t s);
// Image onto the isotropic grid `od`: lo / fr hold nx + ny + nz source index parts and fractions
// (x axis first, then y, then z).
void launch_resample_iso(const float* src, Dims sd, float* dst, Dims od, const int* lo, const float* fr, hipStream_t s);

// --- compose: out = w + u(x + w) on one grid; l maps a displacement to index offsets ---
void launch_compose(const float* displacement, const float* warping, Dims d, const float l[9], float* out, hipStream_t s);
// Two compositions on one grid in one launch.
void launch_compose2(const float* u0, const float* w0, float* out0, const float* u1, const float* w1, float* out1, Dims d,
                     const float l[9], const float* factors, hipStream_t s);

// --- Gaussian smoothing, separable, clamped edges; final pass zeroes faces and blends ---
void launch_gauss_axis(const float* in, float* out, Dims d, int axis, const float* taps_dev, int radius,
                       const float* orig, float w1, float w2, int final_pass, hipStream_t s);
// x and y in one pass into `tmp`, then z (with the blend against `in` and the zeroed faces) into
// `out`; taps per axis, radius at most 8 each (gauss3_fits).
struct GaussTaps3 {
  float x[17], y[17], z[17];
  int rx, ry, rz;
};
bool gauss3_fits(const Dims& d, const GaussTaps3& t);
// With in1 / tmp1 (and tmp1 / in1 / out1 for the z pass) a second field of the same grid is
// smoothed in the same launch.
void launch_gauss_xy(const float* in, float* tmp, Dims d, const GaussTaps3& t, hipStream_t s, const float* in1 = nullptr,
                     float* tmp1 = nullptr);
void launch_gauss_z(const float* tmp, const float* in, float* out, Dims d, const GaussTaps3& t, float w1, float w2,
                    hipStream_t s, const float* tmp1 = nullptr, const float* in1 = nullptr, float* out1 = nullptr);
// All three axes in one launch when the three radii are equal (false, and nothing launched,
// otherwise); the same results as launch_gauss_xy then launch_gauss_z. With in1 / out1 a second
// field of the same grid is smoothed in the same launch.
bool launch_gauss3(const float* in, float* out, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s,
                   const float* in1 = nullptr, float* out1 = nullptr);
// Whether the one-pass form is the faster on this grid: it walks each block's planes in sequence,
// which pays on large grids only (below about four million voxels the two-pass grid, one block per
// plane tile, finishes first).
bool gauss3_pays(const Dims& d);

// --- inversion (ITK InvertDisplacementFieldImageFilter) ---
struct InvertState {
  // written by block 0 of each launch; read by the host after the last launch
  float mean_error;
  float max_error;
  int32_t executed;
  int32_t stop[2];  // stop[i & 1] set: iteration i does not run
  int32_t pad;
};
// What the host reads of one inversion, in coherent host memory: block 0 of the first launch and of
// every executed step stores it, and the host reads it once the inversion's launches are done.
struct InvertReport {
  float mean_error;
  float max_error;
  int32_t executed;
  int32_t pad;
};
// The residual of the initial estimate into `residual` (a field), its partials (2 * blocks floats:
// sums, then maxima) and the state and report reset.
void launch_inv_begin(const float* field, const float* psi, float* residual, Dims d, const float l[9], float3 inv_spacing,
                      InvertState* state, InvertReport* report, float* partials, int blocks, int max_iterations,
                      hipStream_t s);
// One iteration; `partials` holds two such buffers (4 * blocks floats), alternating by step parity.
// The launch that ends the loop stores `tag` to `mailbox` (host memory) as a hint.
void launch_inv_iter(const float* field, float* psi, float* residual, Dims d, const float l[9], float3 inv_spacing,
                     InvertState* state, InvertReport* report, float* partials, int blocks, int step, int max_iterations,
                     double max_tol, double mean_tol, int enforce_boundary, uint32_t* mailbox, uint32_t tag, hipStream_t s);
int inv_blocks(const Dims& d, int cap = 2048);

// --- batches: one launch over several items on one grid (the subjects of a batched SyN level).
// Every item's results are the bits its own launch gives: the items only share the dispatch.
constexpr int kBatchFields = 16;  // fields per smoothing, composition or inversion launch
constexpr int kBatchItems = 8;    // warps, CC pairs, norm pairs or step pairs per launch
constexpr int kBatchChannels = 4; // channels combined per update
// Field pointers of a batched field pass: `in` is read, `orig` is the second input (the blend
// source of a z pass, the warping field of a composition), `out` is written.
struct FieldPtrs {
  const float* in[kBatchFields];
  const float* orig[kBatchFields];
  float* out[kBatchFields];
  int count;
};
struct WarpItems {
  WarpArgs a[kBatchItems];
  int count;
};
void launch_warp_items(const WarpItems& w, hipStream_t s);
struct CcItems {
  CcArgs a[kBatchItems];  // one grid (d, radius, tile_z, probes) for every item; no masks
  int count;
};
void launch_cc_items(const CcItems& c, hipStream_t s);
struct NormItems {
  const float* a[kBatchItems];
  const float* b[kBatchItems];
  float* partial_max[kBatchItems];  // 2 * blocks per item, as launch_field_norms lays them out
  float* partial_sq[kBatchItems];
  int count;
};
void launch_field_norms_items(const NormItems& it, long long n, float3 inv_spacing, int blocks, hipStream_t s);
// Per item: launch_step_factors of its two fields, factors[2 q + f] and maxima[2 q + f].
void launch_step_factors_items(const NormItems& it, int blocks, double step, float* factors, float* maxima, hipStream_t s);
// The channel weighting of a multi-metric update (syn.rs:203-297), per subject and side: the
// magnitude |d_c| = sqrt(sum of its squares, the block partials added in order in f64), the
// coefficient w_c / |d_c| (f64, rounded once) for the channels with |d_c| > eps, and the scale
// (sum of the magnitudes) / channels. coef holds per (subject, side) kBatchChannels coefficients,
// then the scale, then a use mask (as bits); bad[q] is set for a non-finite magnitude of subject q.
struct WeightItems {
  const float* partial_sq[kBatchItems][kBatchChannels];  // the channel's launch_field_norms partials
  double weight[kBatchChannels];
  int channels;
  int count;
};
void launch_channel_weights(const WeightItems& w, int blocks, float* coef, uint32_t* bad, hipStream_t s);
// out[q] = (fma chain over the used channels c of coef_c * d[q][c], from zero) * scale, item q
// reading coef + q * (kBatchChannels + 2).
struct CombineItems {
  const float* d[kBatchFields][kBatchChannels];
  float* out[kBatchFields];
  int channels;
  int count;
};
void launch_combine_items(const CombineItems& c, const float* coef, long long count, hipStream_t s);
// The batched smoothing passes (false, and nothing launched, when the grid would not fit).
bool launch_gauss_xy_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, hipStream_t s);
bool launch_gauss_z_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s);
bool launch_gauss3_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s);
// out = w + u(x + w) with u = in, w = orig; with factors, u is scaled by factors[f] as it is read.
void launch_compose_set(const FieldPtrs& p, Dims d, const float l[9], const float* factors, hipStream_t s);
struct InvItems {
  const float* field[kBatchFields];
  float* psi[kBatchFields];
  float* res[kBatchFields];
  InvertState* state[kBatchFields];
  InvertReport* report[kBatchFields];
  float* partials[kBatchFields];  // 4 * blocks floats each
  int count;
};
// The mailbox word where item q of a batched inversion stores the tag once it has stopped.
constexpr int kInvItemWords = 8;
void launch_inv_begin_items(const InvItems& it, Dims d, const float l[9], float3 inv_spacing, int blocks, int max_iterations,
                            hipStream_t s);
void launch_inv_iter_items(const InvItems& it, Dims d, const float l[9], float3 inv_spacing, int blocks, int step,
                           int max_iterations, double max_tol, double mean_tol, int enforce_boundary, uint32_t* mailbox,
                           uint32_t tag, hipStream_t s);

// --- recursive Gaussian (Deriche, 4th order) along one axis, constant extension at both ends ---
struct DericheCoefficients {
  double n[4], d[4], m[4], bn[4], bm[4];
};
// `in` and `out` must differ; every line needs at least 4 voxels.
// With `lds`, lines along x go through LDS in tiles (the same bits; lines too long for the tile
// budget take the direct form).
void launch_deriche_axis(const float* in, float* out, Dims d, int axis, const DericheCoefficients& c, hipStream_t s,
                         bool lds = true);


}  // namespace fw::hip
