// gfx950 kernels of the SyN hot path: fused CC pair, warp/resample/compose gathers, separable
// Gaussian smoothing, the fixed-point field inversion, and the reductions the SyN loop reads back.
//
// Layout: images are f32[n], fields three f32[n] planes (x, y, z), x-fastest. Every kernel is a
// memory-bound stencil or gather; launches use 64-wide x rows so each wavefront reads and writes
// 256 contiguous bytes per plane. Reductions are fixed trees per block followed by a fixed-order
// combination, so results do not depend on scheduling.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "device_common.hpp"
#include "fastdiv.hpp"

namespace fw::hip {

LaunchEvents*& current_launch_events() {
  static thread_local LaunchEvents* open = nullptr;
  return open;
}

// ------------------------------------------------------------------------------------------
// CC pair (metric/cc.rs:170-321 and wgpu_backend.rs:1354-1520 semantics, fused).
//
// The window sums are those of the f32 voxel values with f64 precision and no rounding history:
// each value is split at a power-of-two step of its image's bound (v = h q + l, h an integer), the
// integer parts are summed exactly in f64 and the remainders in f64 beside them, and the two are
// combined once per window. So a window of zeros sums to zero (as the oracle's direct box sums do,
// which matters for its degenerate-window rule), and a window of tiny values (the edge of a
// zero-padded region, where the local variance is 1e-12 and the coefficient is the largest of the
// image) keeps its relative precision, which a fixed-point quantisation at the bound loses. On
// gfx950 the vector FP64 rate equals the non-packed FP32 rate. The centred moments, the local CC
// and the coefficients follow cc.rs:270-300 in f64; the derivative is the coefficient times the
// f32 central difference, rounded once to f32. Each output voxel writes both derivative
// directions; localCC is symmetric, so one value partial serves both.
//
// A block owns a 64 x 8 (x, y) tile over a run of tile_z planes. Each thread keeps the running
// z-window sums of its columns of the tile and halo in registers: one plane enters and one leaves
// per step, read from global memory (the leaving plane was read 2r+1 steps before and sits in
// cache). The z sums of the step's plane go to LDS; the x pass sums pairs of outputs along x
// (sharing their middle 2r columns, summed directly) and writes its sums back over the z sums after
// a barrier, so one LDS buffer serves both passes; the y pass sums 2r+1 rows per output voxel.
// ------------------------------------------------------------------------------------------
constexpr int kCcTX = 64, kCcTY = 8;

// Threads per block: one per output voxel of the 64 x 8 tile. Each thread owns up to
// ceil((64 + 2R) (8 + 2R) / 512) tile-and-halo columns (three at radius 4); the running sums of a
// column take 20 registers, which leaves one block of eight waves per CU.
template <int R>
constexpr int cc_threads() {
  return kCcTX * kCcTY;
}

// Central differences along the three physical axes on an axis-aligned grid (probes at +-0.5 on
// one grid axis each): the difference of the two linear interpolations is half the difference of
// the two neighbours (the centre cancels), computed as such so that it carries one rounding.
__device__ __forceinline__ void axis_gradient(const float* __restrict__ img, const Dims& d, const Probes& pr, int i, int j,
                                              int k, float out[3]) {
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  const float* p = img + lin(d, i, j, k);
  const int pos[3] = {i, j, k};
  const int len[3] = {d.nx, d.ny, d.nz};
  const long long stride[3] = {1, sy, sz};
  float half_diff[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    // the lower probe at pos - 0.5 clamps to the centre on the first voxel
    const float hi = pos[ax] + 1 < len[ax] ? p[stride[ax]] : 0.0f;
    const float lo = pos[ax] >= 1 ? p[-stride[ax]] : p[0];
    half_diff[ax] = 0.5f * (hi - lo);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ax = pr.o[a][0] != 0.0f ? 0 : (pr.o[a][1] != 0.0f ? 1 : 2);
    const float sgn = pr.o[a][ax];
    const float h = ax == 0 ? half_diff[0] : (ax == 1 ? half_diff[1] : half_diff[2]);
    // the upper probe of the last voxel lies at n - 0.5, outside the buffer
    out[a] = pos[ax] >= len[ax] - 1 ? 0.0f : (sgn > 0.0f ? h : -h) * pr.inv_spacing[a];
  }
}

__device__ float block_sum_f64(double v) {
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
  return (float)t;
}

// The split scale for an image bounded by `bound` (the host's cc_split_scale, on the device).
__device__ __forceinline__ float cc_split_scale_dev(float bound) {
  if (!(bound > 0.0f) || !isfinite(bound)) return 524288.0f;
  int e = 0;
  frexpf(bound, &e);  // bound < 2^e
  return ldexpf(1.0f, 19 - e);
}

// One voxel's values split at the power-of-two step q = 1 / scale: v = h q + l with h an integer
// (|h| <= 2^20 for |v| <= twice the bound) and l exact in f32 (|l| <= q / 2).
struct CcSplit {
  float h, l;
};
__device__ __forceinline__ CcSplit cc_split(float v, float scale, float step) {
  const float h = rintf(v * scale);
  return {h, v - h * step};
}

// The running sums of one column over the window's planes, for both images: the integer parts
// (sums of h, h^2 and h_a h_b, exact in f64: at most 17^3 terms below 2^40) and the corrections of
// the remainders (sums of l, l (2 h q + l) and h_a q l_b + l_a h_b q + l_a l_b), whose rounding is
// relative to their own size (about q). A plane enters and one leaves per step; the integer parts
// return exactly to zero with the values, so a window of zeros has zero sums up to the rounding of
// the corrections (below 1e-19 for a bound of 1).
struct CcColumn {
  double e1a, e1b, e2a, e2b, e2ab;
  double c1a, c1b, c2a, c2b, c2ab;
};

__device__ __forceinline__ void cc_column_step(CcColumn& col, const CcSplit& ea, const CcSplit& eb, const CcSplit& la,
                                               const CcSplit& lb, float qa, float qb) {
  const double hae = ea.h, hbe = eb.h, hal = la.h, hbl = lb.h;
  col.e1a += hae - hal;
  col.e1b += hbe - hbl;
  col.e2a += fma(hae, hae, -hal * hal);
  col.e2b += fma(hbe, hbe, -hbl * hbl);
  col.e2ab += fma(hae, hbe, -hal * hbl);
  const double lae = ea.l, lbe = eb.l, lal = la.l, lbl = lb.l;
  const double Hae = hae * (double)qa, Hbe = hbe * (double)qb, Hal = hal * (double)qa, Hbl = hbl * (double)qb;
  col.c1a += lae - lal;
  col.c1b += lbe - lbl;
  col.c2a += lae * (2.0 * Hae + lae) - lal * (2.0 * Hal + lal);
  col.c2b += lbe * (2.0 * Hbe + lbe) - lbl * (2.0 * Hbl + lbl);
  col.c2ab += fma(Hae, lbe, lae * (Hbe + lbe)) - fma(Hal, lbl, lal * (Hbl + lbl));
}

template <int R, bool MASKED, bool DIAGONAL>
__global__ __launch_bounds__(cc_threads<R>()) void k_cc2(CcArgs a) {
#define FW_CC2_TILE xcd_block()
#include "cc2_body.inc"
#undef FW_CC2_TILE
}

// Several unmasked CC pairs on one grid: the grid's z blocks repeat once per item, and item q's
// blocks are the blocks of its own launch (the same tiles, partials at the same indices of its
// own partial buffers).
template <int R, bool DIAGONAL>
__global__ __launch_bounds__(cc_threads<R>()) void k_cc2_items(CcItems c) {
  constexpr bool MASKED = false;
  const unsigned gz = gridDim.z / (unsigned)c.count;
  const unsigned q = xcd_block().z / gz;
  const CcArgs& a = c.a[q];
#define FW_CC2_TILE make_uint3(xcd_block().x, xcd_block().y, xcd_block().z - q * gz)
#include "cc2_body.inc"
#undef FW_CC2_TILE
}

template <int R, bool M>
static void launch_cc_r(const CcArgs& a, hipStream_t s) {
  if (a.probes.diagonal) launch((k_cc2<R, M, true>), cc_grid(a.d, a.tile_z), dim3(cc_threads<R>()), 0, s, a);
  else launch((k_cc2<R, M, false>), cc_grid(a.d, a.tile_z), dim3(cc_threads<R>()), 0, s, a);
}

template <int R>
static void launch_cc_items_r(const CcItems& c, hipStream_t s) {
  const CcArgs& a = c.a[0];
  dim3 g = cc_grid(a.d, a.tile_z);
  g.z *= (unsigned)c.count;
  if (a.probes.diagonal) launch((k_cc2_items<R, true>), g, dim3(cc_threads<R>()), 0, s, c);
  else launch((k_cc2_items<R, false>), g, dim3(cc_threads<R>()), 0, s, c);
}

void launch_cc_items(const CcItems& c, hipStream_t s) {
  if (c.count < 1 || c.count > kBatchItems) return;
  switch (c.a[0].radius) {
#define FW_CCI_CASE(R)              \
  case R:                           \
    launch_cc_items_r<R>(c, s);     \
    break;
    FW_CCI_CASE(1)
    FW_CCI_CASE(2)
    FW_CCI_CASE(3)
    FW_CCI_CASE(4)
    FW_CCI_CASE(5)
    FW_CCI_CASE(6)
    FW_CCI_CASE(7)
    FW_CCI_CASE(8)
#undef FW_CCI_CASE
    default:
      break;
  }
}

dim3 cc_grid(const Dims& d, int tile_z) {
  return dim3((d.nx + kCcTX - 1) / kCcTX, (d.ny + kCcTY - 1) / kCcTY, (d.nz + tile_z - 1) / tile_z);
}

// The z run of a CC block: the longest of max_tile, max_tile / 2, ..., 1 planes whose grid still
// has min_blocks blocks, so that a small level is not left to a few long-running blocks.
int cc_tile_for(const Dims& d, int max_tile, int min_blocks) {
  int t = std::max(1, max_tile);
  for (; t > 1; t /= 2) {
    const dim3 g = cc_grid(d, t);
    if (size_t(g.x) * g.y * g.z >= size_t(std::max(1, min_blocks))) break;
  }
  return t;
}

void launch_cc_pair(const CcArgs& a, hipStream_t s) {
  const bool m = a.valid != nullptr;
  switch (a.radius) {
#define FW_CC_CASE(R)                         \
  case R:                                     \
    if (m) launch_cc_r<R, true>(a, s);       \
    else launch_cc_r<R, false>(a, s);        \
    break;
    FW_CC_CASE(1)
    FW_CC_CASE(2)
    FW_CC_CASE(3)
    FW_CC_CASE(4)
    FW_CC_CASE(5)
    FW_CC_CASE(6)
    FW_CC_CASE(7)
    FW_CC_CASE(8)
#undef FW_CC_CASE
    default:
      break;
  }
}

// ------------------------------------------------------------------------------------------
// Field norms (field_ops.rs:191-213): per block max |u/spacing| and sum of |u|^2 of up to two
// fields; the host combines the block partials in block order.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void field_norms_block(const float* __restrict__ a, const float* __restrict__ b, long long n,
                                                  float3 inv_sp, float* partial_max, float* partial_sq, unsigned blk,
                                                  unsigned nblk) {
  float ma = 0.0f, mb = 0.0f, sa = 0.0f, ca = 0.0f, sb = 0.0f, cb = 0.0f;
  for (long long i = (long long)blk * blockDim.x + threadIdx.x; i < n; i += (long long)nblk * blockDim.x) {
    {
      const float x = a[i], y = a[i + n], z = a[i + 2 * n];
      const float px = x * inv_sp.x, py = y * inv_sp.y, pz = z * inv_sp.z;
      ma = nan_max(ma, sqrtf(px * px + py * py + pz * pz));
      const float q = x * x + y * y + z * z - ca;
      const float t = sa + q;
      ca = (t - sa) - q;
      sa = t;
    }
    if (b) {
      const float x = b[i], y = b[i + n], z = b[i + 2 * n];
      const float px = x * inv_sp.x, py = y * inv_sp.y, pz = z * inv_sp.z;
      mb = nan_max(mb, sqrtf(px * px + py * py + pz * pz));
      const float q = x * x + y * y + z * z - cb;
      const float t = sb + q;
      cb = (t - sb) - q;
      sb = t;
    }
  }
  const float rma = block_max(ma);
  const float rsa = block_sum(sa);
  const float rmb = block_max(mb);
  const float rsb = block_sum(sb);
  if (threadIdx.x == 0) {
    partial_max[blk] = rma;
    partial_sq[blk] = rsa;
    partial_max[nblk + blk] = rmb;
    partial_sq[nblk + blk] = rsb;
  }
}

__global__ __launch_bounds__(256) void k_field_norms(const float* __restrict__ a, const float* __restrict__ b, long long n,
                                                     float3 inv_sp, float* partial_max, float* partial_sq) {
  field_norms_block(a, b, n, inv_sp, partial_max, partial_sq, blockIdx.x, gridDim.x);
}

// Item q takes blocks [q * blocks, (q + 1) * blocks) and writes the partials of its own launch.
__global__ __launch_bounds__(256) void k_field_norms_items(NormItems it, long long n, float3 inv_sp, int blocks) {
  const unsigned q = blockIdx.x / (unsigned)blocks;
  field_norms_block(it.a[q], it.b[q], n, inv_sp, it.partial_max[q], it.partial_sq[q], blockIdx.x - q * blocks, blocks);
}

void launch_field_norms_items(const NormItems& it, long long n, float3 inv_spacing, int blocks, hipStream_t s) {
  if (it.count < 1 || it.count > kBatchItems) return;
  launch(k_field_norms_items, dim3(blocks * it.count), dim3(256), 0, s, it, n, inv_spacing, blocks);
}

int norm_blocks(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void launch_field_norms(const float* a, const float* b, long long n, float3 inv_spacing, float* partial_max,
                        float* partial_sq, int blocks, hipStream_t s) {
  launch(k_field_norms, dim3(blocks), dim3(256), 0, s, a, b, n, inv_spacing, partial_max, partial_sq);
}

// Update factors of the SyN loop (syn.rs:422-445): for each of `fields` fields, the NaN-propagating
// maximum of its partial maxima and factor = step / max (step when the maximum is not positive),
// divided in f64 and rounded once to f32, as the host computed it. One block.
__device__ __forceinline__ void step_factors_block(const float* __restrict__ partial_max, int blocks, int fields, double step,
                                                   float* factors, float* maxima) {
  for (int f = 0; f < fields; ++f) {
    float m = 0.0f;
    for (int b = threadIdx.x; b < blocks; b += blockDim.x) m = nan_max(m, partial_max[f * blocks + b]);
    const float r = block_max(m);
    if (threadIdx.x == 0) {
      maxima[f] = r;
      factors[f] = r > 0.0f ? (float)(step / (double)r) : (float)step;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_step_factors(const float* __restrict__ partial_max, int blocks, int fields,
                                                      double step, float* factors, float* maxima) {
  step_factors_block(partial_max, blocks, fields, step, factors, maxima);
}

// One block per item.
__global__ __launch_bounds__(256) void k_step_factors_items(NormItems it, int blocks, double step, float* factors,
                                                            float* maxima) {
  const unsigned q = blockIdx.x;
  step_factors_block(it.partial_max[q], blocks, 2, step, factors + 2 * q, maxima + 2 * q);
}

void launch_step_factors_items(const NormItems& it, int blocks, double step, float* factors, float* maxima, hipStream_t s) {
  if (it.count < 1 || it.count > kBatchItems) return;
  launch(k_step_factors_items, dim3(it.count), dim3(256), 0, s, it, blocks, step, factors, maxima);
}

// One thread per subject: the magnitudes and coefficients of both sides, in the order the host
// weighting takes them (channels in order, each magnitude from its partials summed in block order).
__global__ __launch_bounds__(64) void k_channel_weights(WeightItems w, int blocks, float* coef, uint32_t* bad) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= w.count) return;
  uint32_t nonfinite = 0;
  for (int side = 0; side < 2; ++side) {
    float* out = coef + (size_t)(2 * q + side) * (kBatchChannels + 2);
    double total = 0.0;
    uint32_t use = 0;
    for (int c = 0; c < w.channels; ++c) {
      const float* sq = w.partial_sq[q][c] + side * blocks;
      double s2 = 0.0;
      for (int b = 0; b < blocks; ++b) s2 += (double)sq[b];
      const double mag = sqrt(s2);
      if (!isfinite(mag)) nonfinite = 1;
      total += mag;
      out[c] = 0.0f;
      if (mag > (double)kEps64) {
        out[c] = (float)(w.weight[c] / mag);
        use |= 1u << c;
      }
    }
    out[kBatchChannels] = (float)(total / (double)w.channels);
    out[kBatchChannels + 1] = __uint_as_float(use);
  }
  bad[q] = nonfinite;
}

void launch_channel_weights(const WeightItems& w, int blocks, float* coef, uint32_t* bad, hipStream_t s) {
  if (w.count < 1 || w.count > kBatchItems || w.channels < 1 || w.channels > kBatchChannels) return;
  launch(k_channel_weights, dim3(1), dim3(64), 0, s, w, blocks, coef, bad);
}

// Item q takes blocks [q * per, (q + 1) * per): acc = 0, acc = fma(coef_c, d_c, acc) over the
// used channels in order, then acc * scale, as zero_field, axpy per channel and scale compute it.
__global__ __launch_bounds__(256) void k_combine_items(CombineItems c, const float* __restrict__ coef, long long count,
                                                       int per) {
  const unsigned q = blockIdx.x / (unsigned)per;
  const unsigned blk = blockIdx.x - q * per;
  const float* cq = coef + (size_t)q * (kBatchChannels + 2);
  const uint32_t use = __float_as_uint(cq[kBatchChannels + 1]);
  const float scale = cq[kBatchChannels];
  float* out = c.out[q];
  for (long long i = (long long)blk * blockDim.x + threadIdx.x; i < count; i += (long long)per * blockDim.x) {
    float acc = 0.0f;
    for (int ch = 0; ch < c.channels; ++ch)
      if (use & (1u << ch)) acc = fmaf(cq[ch], c.d[q][ch][i], acc);
    out[i] = acc * scale;
  }
}

void launch_step_factors(const float* partial_max, int blocks, int fields, double step, float* factors, float* maxima,
                         hipStream_t s) {
  launch(k_step_factors, dim3(1), dim3(256), 0, s, partial_max, blocks, fields, step, factors, maxima);
}

__global__ __launch_bounds__(256) void k_scale_by(float* __restrict__ x, long long count, const float* __restrict__ factor) {
  const float f = *factor;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    x[i] *= f;
}

// x *= factor[0] and y *= factor[1] in one launch: the first half of the grid takes x.
__global__ __launch_bounds__(256) void k_scale_by2(float* __restrict__ x, float* __restrict__ y, long long count,
                                                   const float* __restrict__ factor) {
  const unsigned half = gridDim.x / 2;
  const bool second = blockIdx.x >= half;
  float* v = second ? y : x;
  const float f = factor[second ? 1 : 0];
  const long long b = second ? blockIdx.x - half : blockIdx.x;
  for (long long i = b * blockDim.x + threadIdx.x; i < count; i += (long long)half * blockDim.x) v[i] *= f;
}


// Finite minimum and maximum of x per block; a block without a finite value reports +inf / -inf.
__global__ __launch_bounds__(256) void k_finite_range(const float* __restrict__ x, long long n, float* partial_min,
                                                      float* partial_max) {
  float lo = INFINITY, hi = -INFINITY;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    if (isfinite(v)) {
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
  }
  __shared__ float red[2][4];
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_down(lo, off, 64));
    hi = fmaxf(hi, __shfl_down(hi, off, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = lo;
    red[1][w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
      lo = fminf(lo, red[0][q]);
      hi = fmaxf(hi, red[1][q]);
    }
    partial_min[blockIdx.x] = lo;
    partial_max[blockIdx.x] = hi;
  }
}

void launch_finite_range(const float* x, long long n, float* partial_min, float* partial_max, int blocks, hipStream_t s) {
  launch(k_finite_range, dim3(blocks), dim3(256), 0, s, x, n, partial_min, partial_max);
}

// ------------------------------------------------------------------------------------------
// Device preprocessing of raw input images (preprocess.rs:40-232 and ITK ImageMomentsCalculator
// as init.rs:47-86 uses it), so that a registration can start from images already in HBM.
// ------------------------------------------------------------------------------------------

static int stream_blocks(long long count);

// Per block: finite minimum, maximum and the count of non-finite voxels.
__global__ __launch_bounds__(256) void k_range_and_bad(const float* __restrict__ x, long long n, float* partial_min,
                                                       float* partial_max, unsigned int* partial_bad) {
  float lo = INFINITY, hi = -INFINITY;
  unsigned int bad = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    if (isfinite(v)) {
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    } else {
      ++bad;
    }
  }
  __shared__ float red[2][4];
  __shared__ unsigned int nb[4];
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_down(lo, off, 64));
    hi = fmaxf(hi, __shfl_down(hi, off, 64));
    bad += __shfl_down(bad, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = lo;
    red[1][w] = hi;
    nb[w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; ++q) {
      lo = fminf(lo, red[0][q]);
      hi = fmaxf(hi, red[1][q]);
      bad += nb[q];
    }
    partial_min[blockIdx.x] = lo;
    partial_max[blockIdx.x] = hi;
    partial_bad[blockIdx.x] = bad;
  }
}

void launch_range_and_bad(const float* x, long long n, float* partial_min, float* partial_max, unsigned int* partial_bad,
                          int blocks, hipStream_t s) {
  launch(k_range_and_bad, dim3(blocks), dim3(256), 0, s, x, n, partial_min, partial_max, partial_bad);
}

// The 256-bin intensity histogram of the finite voxels. A voxel's bin is the number of bin-start
// thresholds (bins 1..255) at or below it: thresholds[k - 1] is the smallest f32 whose f64 value
// is at least the f64 start of bin k, so the f32 comparison equals the f64 one and the bins are
// those of the f64 histogram. Per-block counts go to the block's slot (u64).
__global__ __launch_bounds__(256) void k_intensity_hist(const float* __restrict__ x, long long n,
                                                        const float* __restrict__ thresholds,
                                                        unsigned long long* __restrict__ slots) {
  __shared__ float t[kIntensityBins - 1];
  __shared__ unsigned int h[kIntensityBins];
  for (int i = threadIdx.x; i < kIntensityBins - 1; i += blockDim.x) t[i] = thresholds[i];
  for (int i = threadIdx.x; i < kIntensityBins; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    if (!isfinite(v)) continue;
    int lo = 0, cnt = kIntensityBins - 1;  // count of thresholds <= v, by bisection
    while (cnt > 0) {
      const int half = cnt >> 1;
      if (t[lo + half] <= v) {
        lo += half + 1;
        cnt -= half + 1;
      } else {
        cnt = half;
      }
    }
    atomicAdd(&h[lo], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kIntensityBins; i += blockDim.x)
    slots[(size_t)blockIdx.x * kIntensityBins + i] = (unsigned long long)h[i];
}

void launch_intensity_hist(const float* x, long long n, const float* thresholds, unsigned long long* slots, int blocks,
                           hipStream_t s) {
  launch(k_intensity_hist, dim3(blocks), dim3(256), 0, s, x, n, thresholds, slots);
}

// out = x below lo: out_min; above hi: out_max; else x * scale + shift, with scale and shift given
// as f32 pairs (hi + lo) and the product's rounding error recovered by an fma, so the result is the
// f64 map rounded once to f32 (up to ties within about 2^-46 of a rounding boundary).
__global__ __launch_bounds__(256) void k_window(const float* __restrict__ x, float* __restrict__ out, long long n, float lo,
                                                float hi, float sh, float sl, float th, float tl, float out_min,
                                                float out_max) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    float r;
    if (!isfinite(v)) r = v;
    else if (v < lo) r = out_min;
    else if (v > hi) r = out_max;
    else {
      // no contraction here: an fma formed from p + th would break the two-sum
#pragma clang fp contract(off)
      const float p = v * sh;
      const float pe = fmaf(v, sh, -p);  // p + pe = v * sh exactly
      const float s = p + th;            // two-sum of p and th
      const float bb = s - p;
      const float se = (p - (s - bb)) + (th - bb);
      r = s + (se + (pe + fmaf(v, sl, tl)));
    }
    out[i] = r;
  }
}

void launch_window(const float* x, float* out, long long n, float lo, float hi, float sh, float sl, float th, float tl,
                   float out_min, float out_max, hipStream_t s) {
  launch(k_window, dim3(stream_blocks(n)), dim3(256), 0, s, x, out, n, lo, hi, sh, sl, th, tl, out_min, out_max);
}

// Per block of a (64 x 4)-column, kMomentPlanes-plane chunk: the sums of v, v * i, v * j, v * k over
// its non-zero voxels (index space), compensated f32; the host adds the blocks in f64.
constexpr int kMomentPlanes = 16;
__global__ __launch_bounds__(256) void k_moments(const float* __restrict__ x, Dims d, float* __restrict__ partial) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k0 = blockIdx.z * kMomentPlanes, k1 = min(k0 + kMomentPlanes, d.nz);
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f}, c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (i < d.nx && j < d.ny) {
    for (int k = k0; k < k1; ++k) {
      const float v = x[lin(d, i, j, k)];
      if (v == 0.0f) continue;
      const float terms[4] = {v, v * (float)i, v * (float)j, v * (float)k};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float y = terms[q] - c[q];
        const float t = s[q] + y;
        c[q] = (t - s[q]) - y;
        s[q] = t;
      }
    }
  }
  const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float t = block_sum(s[q]);
    if (threadIdx.x == 0) partial[(size_t)b * 4 + q] = t;
  }
}

dim3 moments_grid(const Dims& d) {
  return dim3((unsigned)((d.nx + 63) / 64), (unsigned)((d.ny + 3) / 4), (unsigned)((d.nz + kMomentPlanes - 1) / kMomentPlanes));
}

void launch_moments(const float* x, Dims d, float* partial, hipStream_t s) {
  launch(k_moments, moments_grid(d), dim3(256), 0, s, x, d, partial);
}

// Image values at continuous indices (mi.rs:161-207 and interp.rs:43-87 rules): the nearest voxel,
// rounding half up, NaN outside [0, n); the ITK linear interpolation, NaN outside [-0.5, n - 0.5).
__global__ __launch_bounds__(256) void k_sample_values(const float* __restrict__ img, Dims d,
                                                       const float4* __restrict__ idx, long long m,
                                                       float* __restrict__ nearest, float* __restrict__ linear) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const float4 c = idx[t];
  const long long q = nearest_index(d, c.x, c.y, c.z);
  nearest[t] = q >= 0 ? img[q] : NAN;
  linear[t] = inside(d, c.x, c.y, c.z) ? tri(img, d, c.x, c.y, c.z) : NAN;
}

void launch_sample_values(const float* img, Dims d, const float4* idx, long long m, float* nearest, float* linear,
                          hipStream_t s) {
  launch(k_sample_values, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, img, d, idx, m, nearest, linear);
}

// ------------------------------------------------------------------------------------------
// Elementwise scale and axpy over the three planes (field_ops.rs:216-244).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scale(float* __restrict__ x, long long count, float s) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    x[i] *= s;
}

__global__ __launch_bounds__(256) void k_axpy(float* __restrict__ acc, const float* __restrict__ x, long long count, float a) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    acc[i] = fmaf(a, x[i], acc[i]);
}

static int stream_blocks(long long count) {
  long long b = (count + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

void launch_scale_by(float* x, long long count, const float* factor, hipStream_t st) {
  launch(k_scale_by, dim3(stream_blocks(count)), dim3(256), 0, st, x, count, factor);
}

void launch_scale_by2(float* x, float* y, long long count, const float* factors, hipStream_t st) {
  launch(k_scale_by2, dim3(2 * stream_blocks(count)), dim3(256), 0, st, x, y, count, factors);
}

void launch_scale(float* x, long long count, float s, hipStream_t st) {
  launch(k_scale, dim3(stream_blocks(count)), dim3(256), 0, st, x, count, s);
}

void launch_axpy(float* acc, const float* x, long long count, float a, hipStream_t st) {
  launch(k_axpy, dim3(stream_blocks(count)), dim3(256), 0, st, acc, x, count, a);
}

void launch_combine_items(const CombineItems& c, const float* coef, long long count, hipStream_t s) {
  if (c.count < 1 || c.count > kBatchFields || c.channels < 1 || c.channels > kBatchChannels) return;
  const int per = stream_blocks(count);
  launch(k_combine_items, dim3(per * c.count), dim3(256), 0, s, c, coef, count, per);
}

// ------------------------------------------------------------------------------------------
// Warp (cpu/mod.rs:60-132, wgpu_backend.rs:2023-2130): output voxel x maps to the source
// continuous index to_src.m * x + to_src.c0 + to_src.l * (u(x) [+ pre(q)]), trilinear inside the
// buffer and zero outside; in mask mode the nearest voxel decides {0, 1}.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void warp_at(const WarpArgs& a, int i, int j, int k) {
  const long long g = lin(a.od, i, j, k);
  const long long n = a.od.n();
  const float ux = a.field[g], uy = a.field[g + n], uz = a.field[g + 2 * n];
  const float fi = (float)i, fj = (float)j, fk = (float)k;
  const IndexMap& m = a.to_src;
  float cx = fmaf(m.m[0], fi, fmaf(m.m[1], fj, fmaf(m.m[2], fk, m.c0[0])));
  float cy = fmaf(m.m[3], fi, fmaf(m.m[4], fj, fmaf(m.m[5], fk, m.c0[1])));
  float cz = fmaf(m.m[6], fi, fmaf(m.m[7], fj, fmaf(m.m[8], fk, m.c0[2])));
  cx += m.l[0] * ux + m.l[1] * uy + m.l[2] * uz;
  cy += m.l[3] * ux + m.l[4] * uy + m.l[5] * uz;
  cz += m.l[6] * ux + m.l[7] * uy + m.l[8] * uz;
  if (a.pre) {
    const IndexMap& p = a.to_pre;
    float px = fmaf(p.m[0], fi, fmaf(p.m[1], fj, fmaf(p.m[2], fk, p.c0[0])));
    float py = fmaf(p.m[3], fi, fmaf(p.m[4], fj, fmaf(p.m[5], fk, p.c0[1])));
    float pz = fmaf(p.m[6], fi, fmaf(p.m[7], fj, fmaf(p.m[8], fk, p.c0[2])));
    px += p.l[0] * ux + p.l[1] * uy + p.l[2] * uz;
    py += p.l[3] * ux + p.l[4] * uy + p.l[5] * uz;
    pz += p.l[6] * ux + p.l[7] * uy + p.l[8] * uz;
    if (inside(a.pd, px, py, pz)) {
      float dx, dy, dz;
      tri3(a.pre, a.pd, px, py, pz, dx, dy, dz);
      cx += a.pre_l[0] * dx + a.pre_l[1] * dy + a.pre_l[2] * dz;
      cy += a.pre_l[3] * dx + a.pre_l[4] * dy + a.pre_l[5] * dz;
      cz += a.pre_l[6] * dx + a.pre_l[7] * dy + a.pre_l[8] * dz;
    }
  }
  float v = 0.0f;
  if (a.nearest) {
    const long long q = nearest_index(a.sd, cx, cy, cz);
    v = (q >= 0 && a.src[q] != 0.0f) ? 1.0f : 0.0f;
    if (a.product) v *= a.product[g];
  } else if (inside(a.sd, cx, cy, cz)) {
    v = tri(a.src, a.sd, cx, cy, cz);
  }
  a.out[g] = v;
}

__global__ __launch_bounds__(256) void k_warp(WarpArgs a) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  if (i >= a.od.nx || j >= a.od.ny) return;
  warp_at(a, i, j, bl.z);
}

// Two warps onto one output grid in one launch: planes [0, nz) of the grid take a0, [nz, 2 nz) a1.
__global__ __launch_bounds__(256) void k_warp2(WarpArgs a0, WarpArgs a1) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  const bool second = (int)bl.z >= a0.od.nz;
  const WarpArgs& a = second ? a1 : a0;
  if (i >= a.od.nx || j >= a.od.ny) return;
  warp_at(a, i, j, second ? (int)bl.z - a0.od.nz : (int)bl.z);
}

// Several warps onto one output grid: planes [q nz, (q + 1) nz) of the grid take item q.
__global__ __launch_bounds__(256) void k_warp_items(WarpItems w) {
  const uint3 bl = xcd_block();
  const int nz = w.a[0].od.nz;
  const int q = (int)bl.z / nz;
  const WarpArgs& a = w.a[q];
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  if (i >= a.od.nx || j >= a.od.ny) return;
  warp_at(a, i, j, (int)bl.z - q * nz);
}

void launch_warp_items(const WarpItems& w, hipStream_t s) {
  if (w.count < 1 || w.count > kBatchItems) return;
  const Dims& od = w.a[0].od;
  const dim3 grid((od.nx + 63) / 64, (od.ny + 3) / 4, od.nz * w.count);
  launch(k_warp_items, grid, dim3(256), 0, s, w);
}

void launch_warp(const WarpArgs& a, hipStream_t s) {
  const dim3 grid((a.od.nx + 63) / 64, (a.od.ny + 3) / 4, a.od.nz);
  launch(k_warp, grid, dim3(256), 0, s, a);
}

void launch_warp2(const WarpArgs& a0, const WarpArgs& a1, hipStream_t s) {
  const dim3 grid((a0.od.nx + 63) / 64, (a0.od.ny + 3) / 4, 2 * a0.od.nz);
  launch(k_warp2, grid, dim3(256), 0, s, a0, a1);
}

// Field resampling onto another grid (field_ops.rs:580-605): linear, zero outside the old buffer.
__global__ __launch_bounds__(256) void k_resample(ResampleArgs a) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  const int k = bl.z;
  if (i >= a.od.nx || j >= a.od.ny) return;
  const long long g = lin(a.od, i, j, k);
  const long long n = a.od.n();
  const float fi = (float)i, fj = (float)j, fk = (float)k;
  const IndexMap& m = a.to_src;
  const float cx = fmaf(m.m[0], fi, fmaf(m.m[1], fj, fmaf(m.m[2], fk, m.c0[0])));
  const float cy = fmaf(m.m[3], fi, fmaf(m.m[4], fj, fmaf(m.m[5], fk, m.c0[1])));
  const float cz = fmaf(m.m[6], fi, fmaf(m.m[7], fj, fmaf(m.m[8], fk, m.c0[2])));
  float vx = 0.0f, vy = 0.0f, vz = 0.0f;
  if (inside(a.sd, cx, cy, cz)) {
    tri3(a.src, a.sd, cx, cy, cz, vx, vy, vz);
  }
  a.out[g] = vx;
  a.out[g + n] = vy;
  a.out[g + 2 * n] = vz;
}

// The pipeline's isotropic resample (register.rs:82-142): output voxel (i, j, k) samples the
// source at per-axis indices whose integer parts (`lo`) and fractions (`fr`) the host computed in
// f64; corners past the last voxel carry zero weight. The weights multiply in axis order and the
// corners add in the reference's order (x bit first), in f32.
__global__ __launch_bounds__(256) void k_resample_iso(const float* __restrict__ src, Dims sd, float* __restrict__ dst, Dims od,
                                                      const int* __restrict__ lo, const float* __restrict__ fr) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  const int k = bl.z;
  if (i >= od.nx || j >= od.ny) return;
  const int b[3] = {lo[i], lo[od.nx + j], lo[od.nx + od.ny + k]};
  const float f[3] = {fr[i], fr[od.nx + j], fr[od.nx + od.ny + k]};
  const int n[3] = {sd.nx, sd.ny, sd.nz};
  float acc = 0.0f;
#pragma unroll
  for (int corner = 0; corner < 8; ++corner) {
    float w = 1.0f;
    int q[3];
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int bit = (corner >> a) & 1;
      q[a] = b[a] + bit;
      inside = inside && q[a] < n[a];
      w *= bit ? f[a] : 1.0f - f[a];
    }
    if (inside && w != 0.0f) acc += w * src[(long long)q[0] + (long long)sd.nx * (q[1] + (long long)sd.ny * q[2])];
  }
  dst[(long long)i + (long long)od.nx * (j + (long long)od.ny * k)] = acc;
}

void launch_resample_iso(const float* src, Dims sd, float* dst, Dims od, const int* lo, const float* fr, hipStream_t s) {
  const dim3 grid((od.nx + 63) / 64, (od.ny + 3) / 4, od.nz);
  launch(k_resample_iso, grid, dim3(256), 0, s, src, sd, dst, od, lo, fr);
}

void launch_resample(const ResampleArgs& a, hipStream_t s) {
  const dim3 grid((a.od.nx + 63) / 64, (a.od.ny + 3) / 4, a.od.nz);
  launch(k_resample, grid, dim3(256), 0, s, a);
}

// ------------------------------------------------------------------------------------------
// Compose (field_ops.rs:251-272, ITK ComposeDisplacementFieldsImageFilter): out = w + u(x + w),
// u interpolated on the shared grid at the continuous index (i, j, k) + l * w, zero outside.
// ------------------------------------------------------------------------------------------

struct L9 {
  float v[9];
};

// Planes [f nz, (f + 1) nz) of the grid compose (in[f], orig[f]) = (u, w) into out[f]. With
// `factors` (device, one per field), u is the update before its scaling: the interpolated update
// is multiplied by the field's factor (the SyN loop's step / max, so that the scaled update is
// never stored).
__global__ __launch_bounds__(256) void k_compose(FieldPtrs p, Dims d, L9 l, const float* __restrict__ factors) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  const int f = (int)bl.z / d.nz;
  const int k = (int)bl.z - f * d.nz;
  const float* __restrict__ u = p.in[f];
  const float* __restrict__ w = p.orig[f];
  float* __restrict__ out = p.out[f];
  if (i >= d.nx || j >= d.ny) return;
  const long long g = lin(d, i, j, k), n = d.n();
  const float wx = w[g], wy = w[g + n], wz = w[g + 2 * n];
  const float cx = (float)i + (l.v[0] * wx + l.v[1] * wy + l.v[2] * wz);
  const float cy = (float)j + (l.v[3] * wx + l.v[4] * wy + l.v[5] * wz);
  const float cz = (float)k + (l.v[6] * wx + l.v[7] * wy + l.v[8] * wz);
  float ox = wx, oy = wy, oz = wz;
  if (inside(d, cx, cy, cz)) {
    float ux, uy, uz;
    tri3(u, d, cx, cy, cz, ux, uy, uz);
    if (factors) {
      const float fa = factors[f];
      ux *= fa;
      uy *= fa;
      uz *= fa;
    }
    ox += ux;
    oy += uy;
    oz += uz;
  }
  out[g] = ox;
  out[g + n] = oy;
  out[g + 2 * n] = oz;
}

void launch_compose_set(const FieldPtrs& p, Dims d, const float l[9], const float* factors, hipStream_t s) {
  if (p.count < 1 || p.count > kBatchFields) return;
  L9 lv;
  for (int q = 0; q < 9; ++q) lv.v[q] = l[q];
  const dim3 grid((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz * p.count);
  launch(k_compose, grid, dim3(256), 0, s, p, d, lv, factors);
}

void launch_compose(const float* displacement, const float* warping, Dims d, const float l[9], float* out, hipStream_t s) {
  FieldPtrs p{};
  p.in[0] = displacement;
  p.orig[0] = warping;
  p.out[0] = out;
  p.count = 1;
  launch_compose_set(p, d, l, nullptr, s);
}

void launch_compose2(const float* u0, const float* w0, float* out0, const float* u1, const float* w1, float* out1, Dims d,
                     const float l[9], const float* factors, hipStream_t s) {
  FieldPtrs p{};
  p.in[0] = u0;
  p.orig[0] = w0;
  p.out[0] = out0;
  p.in[1] = u1;
  p.orig[1] = w1;
  p.out[1] = out1;
  p.count = 2;
  launch_compose_set(p, d, l, factors, s);
}

// ------------------------------------------------------------------------------------------
// Separable Gaussian of a field (field_ops.rs:508-574) in two passes over HBM.
// k_gauss_xy: one block per 64 x 16 (x, y) tile of one plane of one component; the plane tile is
// staged in LDS with its edge-clamped x/y halo, smoothed along x (runs of 4 outputs per thread,
// so each LDS value is read once per run) and then along y (4 rows per thread).
// k_gauss_z: one thread per (x, y) column of one component over a run of planes; the z window
// slides through a ring of 2rz+1 registers, and the final blend w1 * s + w2 * u and the zeroed
// faces are applied on the way out.
// The arithmetic is the three-pass filter's in the same order (x, then y, then z, each an fma
// chain over the taps in order), so the result is bit-identical to it.
// ------------------------------------------------------------------------------------------
constexpr int kG3TX = 64, kG3TY = 16, kG3MaxR = 8;

// Grid z runs over field x component x plane: field f reads p.in[f] and writes p.out[f].
__global__ __launch_bounds__(256) void k_gauss_xy(FieldPtrs p, Dims d, GaussTaps3 t) {
  __shared__ float tile[kG3TY + 2 * kG3MaxR][kG3TX + 2 * kG3MaxR];
  __shared__ float xs[kG3TY + 2 * kG3MaxR][kG3TX];
  const int rx = t.rx, ry = t.ry;
  const int LX = kG3TX + 2 * rx, LY = kG3TY + 2 * ry;
  const uint3 bl = xcd_block();
  const int cz = bl.z / d.nz, z = bl.z - cz * d.nz;
  const int comp = cz % 3, f = cz / 3;
  const int x0 = bl.x * kG3TX, y0 = bl.y * kG3TY;
  const long long n = d.n();
  const float* src = p.in[f] + comp * n + (long long)d.nx * d.ny * z;
  float* dst = p.out[f] + comp * n + (long long)d.nx * d.ny * z;
  const int tid = threadIdx.x;
  for (int e = tid; e < LX * LY; e += 256) {
    const int ly = e / LX, lx = e - ly * LX;
    int gx = x0 - rx + lx, gy = y0 - ry + ly;
    gx = gx < 0 ? 0 : (gx > d.nx - 1 ? d.nx - 1 : gx);
    gy = gy < 0 ? 0 : (gy > d.ny - 1 ? d.ny - 1 : gy);
    tile[ly][lx] = src[gx + d.nx * gy];
  }
  __syncthreads();
  for (int e = tid; e < LY * (kG3TX / 4); e += 256) {
    const int ly = e / (kG3TX / 4), lx = (e - ly * (kG3TX / 4)) * 4;
    float v[4 + 2 * kG3MaxR];
#pragma unroll
    for (int q = 0; q < 4 + 2 * kG3MaxR; ++q)
      if (q < 4 + 2 * rx) v[q] = tile[ly][lx + q];
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
    for (int k = 0; k <= 2 * kG3MaxR; ++k) {
      if (k <= 2 * rx) {
        const float w = t.x[k];
        a0 = fmaf(w, v[k], a0);
        a1 = fmaf(w, v[k + 1], a1);
        a2 = fmaf(w, v[k + 2], a2);
        a3 = fmaf(w, v[k + 3], a3);
      }
    }
    xs[ly][lx] = a0;
    xs[ly][lx + 1] = a1;
    xs[ly][lx + 2] = a2;
    xs[ly][lx + 3] = a3;
  }
  __syncthreads();
  const int tx = tid & 63, ty = (tid >> 6) * 4;
  float v[4 + 2 * kG3MaxR];
#pragma unroll
  for (int q = 0; q < 4 + 2 * kG3MaxR; ++q)
    if (q < 4 + 2 * ry) v[q] = xs[ty + q][tx];
  float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k <= 2 * kG3MaxR; ++k) {
    if (k <= 2 * ry) {
      const float w = t.y[k];
#pragma unroll
      for (int r = 0; r < 4; ++r) b[r] = fmaf(w, v[k + r], b[r]);
    }
  }
  const int x = x0 + tx;
  if (x < d.nx) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = y0 + ty + r;
      if (y < d.ny) dst[x + d.nx * y] = b[r];
    }
  }
}

// The same pass for equal x and y radii R known at compile time (the usual case: isotropic
// variance on an isotropic grid): the tile rows are padded to whole 16-byte vectors, every x run
// reads its 4 + 2R inputs as float4, the loops have no run-time bounds, and the tap chains are
// those of k_gauss_xy, so the results are the same bits.
template <int R>
__global__ __launch_bounds__(256) void k_gauss_xy_r(FieldPtrs p, Dims d, GaussTaps3 t) {
  constexpr int LX = kG3TX + 2 * R, LY = kG3TY + 2 * R;
  constexpr int NV = (4 + 2 * R + 3) / 4;  // float4 reads per x run
  constexpr int LXP0 = (LX + 3) / 4 * 4;
  constexpr int LXP = LXP0 > kG3TX - 4 + 4 * NV ? LXP0 : kG3TX - 4 + 4 * NV;
  __shared__ __align__(16) float tile[LY][LXP];
  __shared__ __align__(16) float xs[LY][kG3TX];
  const uint3 bl = xcd_block();
  const int cz = bl.z / d.nz, z = bl.z - cz * d.nz;
  const int comp = cz % 3, f = cz / 3;
  const int x0 = bl.x * kG3TX, y0 = bl.y * kG3TY;
  const long long n = d.n();
  const float* src = p.in[f] + comp * n + (long long)d.nx * d.ny * z;
  float* dst = p.out[f] + comp * n + (long long)d.nx * d.ny * z;
  const int tid = threadIdx.x;
  for (int ly = tid >> 6; ly < LY; ly += 4) {
    int gy = y0 - R + ly;
    gy = gy < 0 ? 0 : (gy > d.ny - 1 ? d.ny - 1 : gy);
    const float* row = src + (long long)d.nx * gy;
    for (int lx = tid & 63; lx < LX; lx += 64) {
      int gx = x0 - R + lx;
      gx = gx < 0 ? 0 : (gx > d.nx - 1 ? d.nx - 1 : gx);
      tile[ly][lx] = row[gx];
    }
  }
  __syncthreads();
  for (int e = tid; e < LY * (kG3TX / 4); e += 256) {
    const int ly = e >> 4, lx = (e & 15) * 4;
    float v[4 * NV];
    const float4* p = reinterpret_cast<const float4*>(&tile[ly][lx]);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const float4 f = p[q];
      v[4 * q] = f.x;
      v[4 * q + 1] = f.y;
      v[4 * q + 2] = f.z;
      v[4 * q + 3] = f.w;
    }
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
    for (int k = 0; k <= 2 * R; ++k) {
      const float w = t.x[k];
      a0 = fmaf(w, v[k], a0);
      a1 = fmaf(w, v[k + 1], a1);
      a2 = fmaf(w, v[k + 2], a2);
      a3 = fmaf(w, v[k + 3], a3);
    }
    *reinterpret_cast<float4*>(&xs[ly][lx]) = make_float4(a0, a1, a2, a3);
  }
  __syncthreads();
  const int tx = tid & 63, ty = (tid >> 6) * 4;
  float v[4 + 2 * R];
#pragma unroll
  for (int q = 0; q < 4 + 2 * R; ++q) v[q] = xs[ty + q][tx];
  float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k <= 2 * R; ++k) {
    const float w = t.y[k];
#pragma unroll
    for (int r = 0; r < 4; ++r) b[r] = fmaf(w, v[k + r], b[r]);
  }
  const int x = x0 + tx;
  if (x < d.nx) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = y0 + ty + r;
      if (y < d.ny) dst[x + d.nx * y] = b[r];
    }
  }
}

// All three axes in one pass when the three radii are equal (R): a block owns a 64 x 16 (x, y) tile
// of one component over a run of z planes. Each input plane (edge-clamped, with its x/y halo) is
// smoothed along x and y exactly as in k_gauss_xy_r; a thread keeps the x/y results of its four
// outputs for the last 2R + 1 planes in registers and finishes the z pass, the blend and the zeroed
// faces as k_gauss_z does, so the results are the bits of the two-pass form. The next plane's tile
// is fetched into registers while the current one is smoothed. Grid z runs over z chunk x field x
// component: field f reads p.in[f] (also the blend source) and writes p.out[f].
template <int R>
__global__ __launch_bounds__(256) void k_gauss3(FieldPtrs p, Dims d, GaussTaps3 t, int zrun, float w1, float w2) {
  constexpr int W = 2 * R + 1;
  constexpr int LX = kG3TX + 2 * R, LY = kG3TY + 2 * R;
  constexpr int NV = (4 + 2 * R + 3) / 4;
  constexpr int LXP0 = (LX + 3) / 4 * 4;
  constexpr int LXP = LXP0 > kG3TX - 4 + 4 * NV ? LXP0 : kG3TX - 4 + 4 * NV;
  constexpr int PER = (LX * LY + 255) / 256;  // tile elements per thread
  __shared__ __align__(16) float tile[LY][LXP];
  __shared__ __align__(16) float xs[LY][kG3TX];
  const uint3 bl = xcd_block();
  const int fields = p.count;
  const int cz = bl.z % (3 * fields), chunk = bl.z / (3 * fields);
  const int comp = cz % 3, f = cz / 3;
  const long long n = d.n(), plane = (long long)d.nx * d.ny;
  const float* src = p.in[f] + comp * n;
  float* dst = p.out[f] + comp * n;
  const int x0 = bl.x * kG3TX, y0 = bl.y * kG3TY;
  const int z0 = chunk * zrun, z1 = min(z0 + zrun, d.nz);
  const int tid = threadIdx.x;
  // this thread's tile elements: clamped offset within a plane, -1 past the tile
  int toff[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q;
    toff[q] = -1;
    if (e < LX * LY) {
      const int ly = e / LX, lx = e - ly * LX;
      int gx = x0 - R + lx, gy = y0 - R + ly;
      gx = gx < 0 ? 0 : (gx > d.nx - 1 ? d.nx - 1 : gx);
      gy = gy < 0 ? 0 : (gy > d.ny - 1 ? d.ny - 1 : gy);
      toff[q] = gx + d.nx * gy;
    }
  }
  float pre[PER];
  auto fetch = [&](int zin) {
    const float* p = src + plane * (zin < 0 ? 0 : (zin > d.nz - 1 ? d.nz - 1 : zin));
#pragma unroll
    for (int q = 0; q < PER; ++q) pre[q] = toff[q] >= 0 ? p[toff[q]] : 0.0f;
  };
  const int tx = tid & 63, ty = (tid >> 6) * 4;
  const int x = x0 + tx;
  float ring[W][4];
#pragma unroll
  for (int q = 0; q < W; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) ring[q][r] = 0.0f;
  fetch(z0 - R);
  for (int zin = z0 - R; zin < z1 + R; ++zin) {
#pragma unroll
    for (int q = 0; q < PER; ++q)
      if (toff[q] >= 0) {
        const int e = tid + 256 * q;
        const int ly = e / LX, lx = e - ly * LX;
        tile[ly][lx] = pre[q];
      }
    __syncthreads();
    if (zin + 1 < z1 + R) fetch(zin + 1);
    for (int e = tid; e < LY * (kG3TX / 4); e += 256) {
      const int ly = e >> 4, lx = (e & 15) * 4;
      float v[4 * NV];
      const float4* p = reinterpret_cast<const float4*>(&tile[ly][lx]);
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const float4 f = p[q];
        v[4 * q] = f.x;
        v[4 * q + 1] = f.y;
        v[4 * q + 2] = f.z;
        v[4 * q + 3] = f.w;
      }
      float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
      for (int k = 0; k <= 2 * R; ++k) {
        const float w = t.x[k];
        a0 = fmaf(w, v[k], a0);
        a1 = fmaf(w, v[k + 1], a1);
        a2 = fmaf(w, v[k + 2], a2);
        a3 = fmaf(w, v[k + 3], a3);
      }
      *reinterpret_cast<float4*>(&xs[ly][lx]) = make_float4(a0, a1, a2, a3);
    }
    __syncthreads();
    float v[4 + 2 * R];
#pragma unroll
    for (int q = 0; q < 4 + 2 * R; ++q) v[q] = xs[ty + q][tx];
    float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k <= 2 * R; ++k) {
      const float w = t.y[k];
#pragma unroll
      for (int r = 0; r < 4; ++r) b[r] = fmaf(w, v[k + r], b[r]);
    }
#pragma unroll
    for (int q = 0; q < W - 1; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) ring[q][r] = ring[q + 1][r];
#pragma unroll
    for (int r = 0; r < 4; ++r) ring[W - 1][r] = b[r];
    const int k = zin - R;
    if (k >= z0 && x < d.nx) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int y = y0 + ty + r;
        if (y < d.ny) {
          const long long g = plane * k + x + (long long)d.nx * y;
          float o = 0.0f;
          if (!(x == 0 || y == 0 || x == d.nx - 1 || y == d.ny - 1 || k == 0 || k == d.nz - 1)) {
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < W; ++q) acc = fmaf(t.z[q], ring[q][r], acc);
            o = w2 != 0.0f ? acc * w1 + src[g] * w2 : acc * w1;
          }
          dst[g] = o;
        }
      }
    }
  }
}

// z runs of the one-pass form: about kG3Blocks blocks, each run at least 4R planes long (a run
// smooths 2R planes beyond its own along x and y).
#ifndef FW_G3_BLOCKS
#define FW_G3_BLOCKS 2048
#endif
constexpr int kG3Blocks = FW_G3_BLOCKS;
static int g3_zrun(const Dims& d, int r, int fields) {
  const long long cols = (long long)((d.nx + kG3TX - 1) / kG3TX) * ((d.ny + kG3TY - 1) / kG3TY) * 3 * fields;
  const long long chunks = (kG3Blocks + cols - 1) / cols;
  int run = (int)((d.nz + chunks - 1) / chunks);
  const int least = 4 * r;
  return run < least ? least : run;
}

template <int RZ>
// Grid z runs over z chunk x field x component: field f reads p.in[f] (the x/y pass's result) and
// p.orig[f] (the blend source) and writes p.out[f].
__global__ __launch_bounds__(256) void k_gauss_z(FieldPtrs p, Dims d, GaussTaps3 t, int zrun, float w1, float w2) {
  constexpr int W = 2 * RZ + 1;
  const uint3 bl = xcd_block();
  const int fields = p.count;
  const int cz = bl.z % (3 * fields), chunk = bl.z / (3 * fields);
  const int comp = cz % 3, f = cz / 3;
  const float* __restrict__ in = p.in[f];
  const float* __restrict__ orig = p.orig[f];
  float* __restrict__ out = p.out[f];
  const int x = bl.x * 64 + (threadIdx.x & 63);
  const int y = bl.y * 4 + (threadIdx.x >> 6);
  if (x >= d.nx || y >= d.ny) return;
  const int z0 = chunk * zrun, z1 = min(z0 + zrun, d.nz);
  const long long n = d.n(), plane = (long long)d.nx * d.ny;
  const long long col = (long long)x + (long long)d.nx * y;
  const float* src = in + comp * n + col;
  const float* org = orig + comp * n + col;
  float* dst = out + comp * n + col;
  const bool edge_xy = x == 0 || y == 0 || x == d.nx - 1 || y == d.ny - 1;
  float ring[W];
#pragma unroll
  for (int q = 0; q < W; ++q) ring[q] = 0.0f;
  auto plane_of = [&](int zin) { return zin < 0 ? 0 : (zin > d.nz - 1 ? d.nz - 1 : zin); };
  float next = src[plane * plane_of(z0 - RZ)];
  for (int zin = z0 - RZ; zin < z1 + RZ; ++zin) {
    const float cur = next;
    if (zin + 1 < z1 + RZ) next = src[plane * plane_of(zin + 1)];
#pragma unroll
    for (int q = 0; q < W - 1; ++q) ring[q] = ring[q + 1];
    ring[W - 1] = cur;
    const int k = zin - RZ;
    if (k >= z0) {
      float o = 0.0f;
      if (!(edge_xy || k == 0 || k == d.nz - 1)) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < W; ++q) acc = fmaf(t.z[q], ring[q], acc);
        o = w2 != 0.0f ? acc * w1 + org[plane * k] * w2 : acc * w1;
      }
      dst[plane * k] = o;
    }
  }
}

// The z-run per block: about 2048 blocks, runs at least 4 * rz planes long, so the 2 * rz halo
// planes each run re-reads stay a minority of its input.
static int gz_zrun(const Dims& d, int rz) {
  const long long cols = (long long)((d.nx + 63) / 64) * ((d.ny + 3) / 4) * 3;
  const long long chunks = (2048 + cols - 1) / cols;
  int run = (int)((d.nz + chunks - 1) / chunks);
  const int least = 4 * rz > 8 ? 4 * rz : 8;
  return run < least ? least : run;
}

bool gauss3_fits(const Dims& d, const GaussTaps3& t) {
  return t.rx <= kG3MaxR && t.ry <= kG3MaxR && t.rz <= kG3MaxR && t.rz >= 1 && 6LL * d.nz <= 65535;
}

static FieldPtrs field_ptrs(const float* in0, const float* orig0, float* out0, const float* in1, const float* orig1,
                            float* out1) {
  FieldPtrs p{};
  p.in[0] = in0;
  p.orig[0] = orig0;
  p.out[0] = out0;
  p.count = 1;
  if (in1) {
    p.in[1] = in1;
    p.orig[1] = orig1;
    p.out[1] = out1;
    p.count = 2;
  }
  return p;
}

bool launch_gauss_xy_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, hipStream_t s) {
  if (p.count < 1 || p.count > kBatchFields || 3LL * p.count * d.nz > 65535) return false;
  const dim3 gxy((d.nx + kG3TX - 1) / kG3TX, (d.ny + kG3TY - 1) / kG3TY, 3 * p.count * d.nz);
  if (t.rx == t.ry) {
    switch (t.rx) {
#define FW_GXY_CASE(R)                                        \
  case R:                                                     \
    launch(k_gauss_xy_r<R>, gxy, dim3(256), 0, s, p, d, t);  \
    return true;
      FW_GXY_CASE(1)
      FW_GXY_CASE(2)
      FW_GXY_CASE(3)
      FW_GXY_CASE(4)
      FW_GXY_CASE(5)
      FW_GXY_CASE(6)
      FW_GXY_CASE(7)
      FW_GXY_CASE(8)
#undef FW_GXY_CASE
      default:
        break;
    }
  }
  launch(k_gauss_xy, gxy, dim3(256), 0, s, p, d, t);
  return true;
}

void launch_gauss_xy(const float* in, float* tmp, Dims d, const GaussTaps3& t, hipStream_t s, const float* in1,
                     float* tmp1) {
  launch_gauss_xy_set(field_ptrs(in, nullptr, tmp, in1, nullptr, tmp1), d, t, s);
}

bool gauss3_pays(const Dims& d) { return d.n() >= 4000000; }

bool launch_gauss3_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s) {
  if (t.rx != t.ry || t.ry != t.rz) return false;
  if (p.count < 1 || p.count > kBatchFields) return false;
  const int run = g3_zrun(d, t.rz, p.count);
  const dim3 g((d.nx + kG3TX - 1) / kG3TX, (d.ny + kG3TY - 1) / kG3TY, 3 * p.count * ((d.nz + run - 1) / run));
  if (g.z > 65535) return false;
  switch (t.rz) {
#define FW_G3_CASE(R)                                                  \
  case R:                                                              \
    launch(k_gauss3<R>, g, dim3(256), 0, s, p, d, t, run, w1, w2);    \
    return true;
    FW_G3_CASE(1)
    FW_G3_CASE(2)
    FW_G3_CASE(3)
    FW_G3_CASE(4)
    FW_G3_CASE(5)
    FW_G3_CASE(6)
    FW_G3_CASE(7)
    FW_G3_CASE(8)
#undef FW_G3_CASE
    default:
      return false;
  }
}

bool launch_gauss3(const float* in, float* out, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s,
                   const float* in1, float* out1) {
  return launch_gauss3_set(field_ptrs(in, nullptr, out, in1, nullptr, out1), d, t, w1, w2, s);
}

bool launch_gauss_z_set(const FieldPtrs& p, Dims d, const GaussTaps3& t, float w1, float w2, hipStream_t s) {
  const int run = gz_zrun(d, t.rz);
  const dim3 gz((d.nx + 63) / 64, (d.ny + 3) / 4, 3 * p.count * ((d.nz + run - 1) / run));
  if (p.count < 1 || p.count > kBatchFields || gz.z > 65535) return false;
  switch (t.rz) {
#define FW_GZ_CASE(R)                                                  \
  case R:                                                              \
    launch(k_gauss_z<R>, gz, dim3(256), 0, s, p, d, t, run, w1, w2);  \
    return true;
    FW_GZ_CASE(1)
    FW_GZ_CASE(2)
    FW_GZ_CASE(3)
    FW_GZ_CASE(4)
    FW_GZ_CASE(5)
    FW_GZ_CASE(6)
    FW_GZ_CASE(7)
    FW_GZ_CASE(8)
#undef FW_GZ_CASE
    default:
      return false;
  }
}

void launch_gauss_z(const float* tmp, const float* in, float* out, Dims d, const GaussTaps3& t, float w1, float w2,
                    hipStream_t s, const float* tmp1, const float* in1, float* out1) {
  launch_gauss_z_set(field_ptrs(tmp, in, out, tmp1, in1, out1), d, t, w1, w2, s);
}

// ------------------------------------------------------------------------------------------
// Recursive Gaussian (pyramid.rs:184-305, ITK RecursiveGaussianImageFilter, zero order): one
// thread per line along `axis`. The causal pass writes its result to `out`; the anti-causal pass
// walks back, adds its own result and overwrites `out`. Lines of neighbouring threads are
// adjacent in x for the y and z axes (coalesced); x lines stay in the same cache lines for 16
// consecutive voxels.
// ------------------------------------------------------------------------------------------
// The recursions of one line: input p, output q, element i at offset i * stride (global memory,
// or a line staged in LDS).
__device__ __forceinline__ void deriche_line(const float* __restrict__ p, float* __restrict__ q, long long stride, int len,
                                             const DericheCoefficients& c) {
  // The recursions run in f64 (the vector FP64 rate of gfx950 is that of FP32): in f32 their
  // rounding grows with sigma as the poles approach 1. The causal part is stored in f32 and the
  // anti-causal part added to it, so the result carries two f32 roundings.
  // causal: y[i] = sum n_k x[i-k] - sum d_k y[i-1-k], with x and y extended by x[0] at the start
  {
    const double v1 = p[0];
    double x0 = v1, x1 = v1, x2 = v1, x3 = v1;  // x[i], x[i-1], x[i-2], x[i-3]
    double y1 = 0.0, y2 = 0.0, y3 = 0.0, y4 = 0.0;
    for (int i = 0; i < len; ++i) {
      x3 = x2;
      x2 = x1;
      x1 = x0;
      x0 = p[(long long)i * stride];
      double y = x0 * c.n[0] + x1 * c.n[1] + x2 * c.n[2] + x3 * c.n[3];
      // outputs before the start of the line are replaced by the boundary terms v1 * bn_k
      const double sub = (i >= 1 ? y1 * c.d[0] : v1 * c.bn[0]) + (i >= 2 ? y2 * c.d[1] : v1 * c.bn[1]) +
                         (i >= 3 ? y3 * c.d[2] : v1 * c.bn[2]) + (i >= 4 ? y4 * c.d[3] : v1 * c.bn[3]);
      y -= sub;
      y4 = y3;
      y3 = y2;
      y2 = y1;
      y1 = y;
      q[(long long)i * stride] = (float)y;
    }
  }
  // anti-causal: a[i-1] = sum m_k x[i+k] - sum d_k a[i+k], with x extended by x[len-1] at the end
  {
    const double v2 = p[(long long)(len - 1) * stride];
    double x0 = v2, x1 = v2, x2 = v2, x3 = v2;  // x[i], x[i+1], x[i+2], x[i+3] for a[i-1]
    double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;  // a[i], a[i+1], a[i+2], a[i+3]
    for (int t = 0; t < len; ++t) {
      const int i = len - t;  // this step produces a[i-1]
      // x[i..i+3]: x[i] is the next input inward, past the end it is v2
      x3 = x2;
      x2 = x1;
      x1 = x0;
      x0 = i <= len - 1 ? (double)p[(long long)i * stride] : v2;
      double y = x0 * c.m[0] + x1 * c.m[1] + x2 * c.m[2] + x3 * c.m[3];
      const double sub = (t >= 1 ? a1 * c.d[0] : v2 * c.bm[0]) + (t >= 2 ? a2 * c.d[1] : v2 * c.bm[1]) +
                         (t >= 3 ? a3 * c.d[2] : v2 * c.bm[2]) + (t >= 4 ? a4 * c.d[3] : v2 * c.bm[3]);
      y -= sub;
      a4 = a3;
      a3 = a2;
      a2 = a1;
      a1 = y;
      const long long o = (long long)(i - 1) * stride;
      q[o] = (float)((double)q[o] + y);
    }
  }
}

__global__ __launch_bounds__(256) void k_deriche_axis(const float* __restrict__ in, float* __restrict__ out, Dims d,
                                                      int axis, DericheCoefficients c) {
  const long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)d.nx * d.ny;
  long long lines, base, stride;
  int len;
  if (axis == 0) {
    lines = (long long)d.ny * d.nz;
    base = line * d.nx;
    stride = 1;
    len = d.nx;
  } else if (axis == 1) {
    lines = (long long)d.nx * d.nz;
    base = (line % d.nx) + (line / d.nx) * plane;
    stride = d.nx;
    len = d.ny;
  } else {
    lines = plane;
    base = line;
    stride = plane;
    len = d.nz;
  }
  if (line >= lines) return;
  deriche_line(in + base, out + base, stride, len, c);
}

// Lines along x (contiguous in memory, so one thread per line reads nothing coalesced): a block
// of one wave stages `per` consecutive lines in LDS with coalesced loads, runs one line's
// recursions per thread on LDS, and stores the outputs coalesced. The same recursions as
// k_deriche_axis, so the same bits.
constexpr int kDericheLdsBytes = 48 << 10;
__global__ __launch_bounds__(64) void k_deriche_x(const float* __restrict__ in, float* __restrict__ out, Dims d, int per,
                                                  DericheCoefficients c) {
  extern __shared__ float tile[];
  const int nx = d.nx, row = nx + 1;
  const long long lines = (long long)d.ny * d.nz;
  const long long first = (long long)blockIdx.x * per;
  const int nl = (int)(lines - first < per ? lines - first : per);
  float* tin = tile;
  float* tout = tile + per * row;
  const float* src = in + first * nx;
  for (int l = 0; l < nl; ++l)
    for (int x = threadIdx.x; x < nx; x += blockDim.x) tin[l * row + x] = src[(long long)l * nx + x];
  __syncthreads();
  if ((int)threadIdx.x < nl) deriche_line(tin + threadIdx.x * row, tout + threadIdx.x * row, 1, nx, c);
  __syncthreads();
  float* dst = out + first * nx;
  for (int l = 0; l < nl; ++l)
    for (int x = threadIdx.x; x < nx; x += blockDim.x) dst[(long long)l * nx + x] = tout[l * row + x];
}

void launch_deriche_axis(const float* in, float* out, Dims d, int axis, const DericheCoefficients& c, hipStream_t s,
                         bool lds) {
  const long long lines = axis == 0 ? (long long)d.ny * d.nz : axis == 1 ? (long long)d.nx * d.nz : (long long)d.nx * d.ny;
  if (axis == 0 && lds) {
    const int per = (int)(kDericheLdsBytes / (2 * sizeof(float) * (size_t)(d.nx + 1)));
    if (per >= 4) {
      const int p = per > 64 ? 64 : per;
      const size_t bytes = 2 * sizeof(float) * (size_t)p * (size_t)(d.nx + 1);
      launch(k_deriche_x, dim3((unsigned)((lines + p - 1) / p)), dim3(64), (uint32_t)bytes, s, in, out, d, p, c);
      return;
    }
  }
  launch(k_deriche_axis, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s, in, out, d, axis, c);
}

// ------------------------------------------------------------------------------------------
// Separable Gaussian (field_ops.rs:508-574): taps clamped at the edges; the final pass zeroes
// every face voxel and blends w1 * smooth + w2 * original.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gauss_axis(const float* __restrict__ in, float* __restrict__ out, Dims d, int axis,
                                                    const float* __restrict__ taps, int radius,
                                                    const float* __restrict__ orig, float w1, float w2, int final_pass) {
  const uint3 bl = xcd_block();
  const int i = bl.x * 64 + (threadIdx.x & 63);
  const int j = bl.y * 4 + (threadIdx.x >> 6);
  const int k = bl.z;
  if (i >= d.nx || j >= d.ny) return;
  const long long g = lin(d, i, j, k), n = d.n();
  if (final_pass && (i == 0 || j == 0 || k == 0 || i == d.nx - 1 || j == d.ny - 1 || k == d.nz - 1)) {
    out[g] = 0.0f;
    out[g + n] = 0.0f;
    out[g + 2 * n] = 0.0f;
    return;
  }
  const int pos = axis == 0 ? i : axis == 1 ? j : k;
  const int len = axis == 0 ? d.nx : axis == 1 ? d.ny : d.nz;
  const long long stride = axis == 0 ? 1 : axis == 1 ? (long long)d.nx : (long long)d.nx * d.ny;
  const long long rowbase = g - (long long)pos * stride;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (int t = 0; t <= 2 * radius; ++t) {
    int q = pos + t - radius;
    q = q < 0 ? 0 : (q > len - 1 ? len - 1 : q);
    const long long src = rowbase + (long long)q * stride;
    const float wt = taps[t];
    ax = fmaf(wt, in[src], ax);
    ay = fmaf(wt, in[src + n], ay);
    az = fmaf(wt, in[src + 2 * n], az);
  }
  if (final_pass) {
    ax = ax * w1 + orig[g] * w2;
    ay = ay * w1 + orig[g + n] * w2;
    az = az * w1 + orig[g + 2 * n] * w2;
  }
  out[g] = ax;
  out[g + n] = ay;
  out[g + 2 * n] = az;
}

void launch_gauss_axis(const float* in, float* out, Dims d, int axis, const float* taps_dev, int radius,
                       const float* orig, float w1, float w2, int final_pass, hipStream_t s) {
  const dim3 grid((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz);
  launch(k_gauss_axis, grid, dim3(256), 0, s, in, out, d, axis, taps_dev, radius, orig, w1, w2, final_pass);
}

// ------------------------------------------------------------------------------------------
// Inversion (field_ops.rs:315-392, ITK InvertDisplacementFieldImageFilter): one launch per
// iteration. The residual e_i = psi_i + field(x + psi_i) is kept in a field buffer: the launch
// for iteration i reads it, applies psi_{i+1} = psi_i + eps * clip(-e_i, eps * max_i), and, if
// another iteration may follow, gathers the field once at x + psi_{i+1} to store e_{i+1} and
// write per-block partials of its scaled norms. The residual at x gathers only the forward field,
// never psi at other voxels, so psi and e are updated in place.
// Iteration i runs only if state->stop[i & 1] is clear. Every block first reduces the partials
// that the previous launch left (the same fixed-order reduction in every block, so all blocks
// agree on mean_i and max_i without a grid barrier or a device-scope fence; a release at device
// scope would write back the whole L2 of each XCD). Block 0 publishes the statistics and
// stop[(i + 1) & 1], a word no block of the same launch reads. Each thread takes kInvVoxels voxels
// per step of its grid-stride loop so their gathers are in flight together.
// ------------------------------------------------------------------------------------------
constexpr int kInvThreads = 256;
constexpr int kInvVoxels = 4;

// The row and plane divisors of a grid's linear index (grids below 2^31 voxels, which the
// inversion's caller checks).
struct InvIndex {
  FastDiv x, y;
};
static InvIndex inv_index(const Dims& d) { return InvIndex{make_fastdiv((uint32_t)d.nx), make_fastdiv((uint32_t)d.ny)}; }

__device__ __forceinline__ void inv_residual_at(const float* __restrict__ field, const Dims& d, const float* l, int i,
                                                int j, int k, float wx, float wy, float wz, float& ex, float& ey,
                                                float& ez) {
  const float cx = (float)i + (l[0] * wx + l[1] * wy + l[2] * wz);
  const float cy = (float)j + (l[3] * wx + l[4] * wy + l[5] * wz);
  const float cz = (float)k + (l[6] * wx + l[7] * wy + l[8] * wz);
  ex = wx;
  ey = wy;
  ez = wz;
  if (inside(d, cx, cy, cz)) {
    float ux, uy, uz;
    tri3(field, d, cx, cy, cz, ux, uy, uz);
    ex += ux;
    ey += uy;
    ez += uz;
  }
}

__device__ __forceinline__ float scaled_norm(float3 inv_sp, float ex, float ey, float ez) {
  const float px = ex * inv_sp.x, py = ey * inv_sp.y, pz = ez * inv_sp.z;
  return sqrtf(px * px + py * py + pz * pz);
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

__device__ __forceinline__ void inv_write_partials(float s, float m, float* partials, int blocks, unsigned chunk) {
  const float bs = block_sum(s);
  const float bm = block_max(m);
  if (threadIdx.x == 0) {
    partials[chunk] = bs;
    partials[blocks + chunk] = bm;
  }
}

// The voxels of block b: the contiguous range of its logical chunk (xcd_slab), in steps of
// kInvThreads * kInvVoxels; chunk lengths are multiples of kInvThreads.
struct InvRange {
  unsigned chunk;
  long long lo, hi;
};
__device__ __forceinline__ InvRange inv_range(long long n, unsigned blk, unsigned nblk) {
  const long long len = ((n + nblk - 1) / nblk + kInvThreads - 1) / kInvThreads * kInvThreads;
  InvRange r;
  r.chunk = xcd_slab(blk, nblk);
  r.lo = (long long)r.chunk * len;
  r.hi = r.lo + len < n ? r.lo + len : n;
  return r;
}

// (i, j, k) of voxel g < 2^31 by two multiply-high divisions (nx, ny in `ix`).
__device__ __forceinline__ void inv_coords(const InvIndex& ix, long long g, int& i, int& j, int& k) {
  const uint32_t u = (uint32_t)g;
  const uint32_t r = fast_div(u, ix.x);
  i = (int)(u - r * ix.x.d);
  const uint32_t q = fast_div(r, ix.y);
  j = (int)(r - q * ix.y.d);
  k = (int)q;
}

__device__ __forceinline__ void inv_begin_block(const float* __restrict__ field, const float* __restrict__ psi,
                                                float* __restrict__ res, const Dims& d, const InvIndex& ix, const L9& l,
                                                float3 inv_sp, InvertState* state, InvertReport* report, float* partials,
                                                int max_iterations, unsigned blk, unsigned nblk) {
  const long long n = d.n();
  const InvRange rg = inv_range(n, blk, nblk);
  float s = 0.0f, c = 0.0f, m = 0.0f;
  for (long long base = rg.lo + threadIdx.x; base < rg.hi; base += kInvThreads * kInvVoxels) {
#pragma unroll
    for (int v = 0; v < kInvVoxels; ++v) {
      const long long g = base + v * kInvThreads;
      if (g < rg.hi) {
        int i, j, k;
        inv_coords(ix, g, i, j, k);
        float ex, ey, ez;
        inv_residual_at(field, d, l.v, i, j, k, psi[g], psi[g + n], psi[g + 2 * n], ex, ey, ez);
        res[g] = ex;
        res[g + n] = ey;
        res[g + 2 * n] = ez;
        const float nrm = scaled_norm(inv_sp, ex, ey, ez);
        m = nan_max(m, nrm);
        kahan_add(s, c, nrm);
      }
    }
  }
  inv_write_partials(s, m, partials, nblk, rg.chunk);
  if (blk == 0 && threadIdx.x == 0) {
    state->mean_error = 0.0f;
    state->max_error = 0.0f;
    state->executed = 0;
    state->stop[0] = max_iterations > 0 ? 0 : 1;
    state->stop[1] = 0;
    report->mean_error = 0.0f;
    report->max_error = 0.0f;
    report->executed = 0;
  }
}

// One block of iteration `step`; `stop_word` receives the tag once the iteration loop ends.
__device__ __forceinline__ void inv_iter_block(const float* __restrict__ field, float* __restrict__ psi,
                                               float* __restrict__ res, const Dims& d, const InvIndex& ix, const L9& l,
                                               float3 inv_sp, InvertState* state, InvertReport* report, float* partials,
                                               int step, int max_iterations, float max_tol, float mean_tol, int enforce,
                                               uint32_t* mailbox, uint32_t* stop_word, uint32_t tag, unsigned blk,
                                               unsigned nblk) {
  if (state->stop[step & 1]) {
    // once stopped, stay stopped: the next launch reads the other word
    if (blk == 0 && threadIdx.x == 0) state->stop[(step + 1) & 1] = 1;
    return;
  }
  const int blocks = (int)nblk;
  // progress hint for the host: this inversion has started iteration `step`
  if (blk == 0 && threadIdx.x == 0)
    __hip_atomic_store(mailbox + 1, ((tag & 0xffffffu) << 8) | (uint32_t)(step + 1), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  const float* in = partials + (step & 1) * 2 * blocks;
  float* out = partials + ((step + 1) & 1) * 2 * blocks;
  float s = 0.0f, c = 0.0f, m = 0.0f;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    kahan_add(s, c, in[b]);
    m = nan_max(m, in[blocks + b]);
  }
  __shared__ float stats[2];
  {
    const float bs = block_sum(s);
    const float bm = block_max(m);
    if (threadIdx.x == 0) {
      stats[0] = bs;
      stats[1] = bm;
    }
    __syncthreads();
  }
  const long long n = d.n();
  const float mean = stats[0] / (float)n;
  const float mx = stats[1];
  const bool next = (step + 1) < max_iterations && mx > max_tol && mean > mean_tol;
  if (blk == 0 && threadIdx.x == 0) {
    state->mean_error = mean;
    state->max_error = mx;
    state->executed = step + 1;
    state->stop[(step + 1) & 1] = next ? 0 : 1;
    report->mean_error = mean;
    report->max_error = mx;
    report->executed = step + 1;
    // a hint for the host: later launches of this inversion would return at once
    if (!next) __hip_atomic_store(stop_word, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  const float eps = step == 0 ? 0.75f : 0.5f;
  const float limit = eps * mx;
  const InvRange rg = inv_range(n, blk, nblk);
  s = 0.0f;
  c = 0.0f;
  m = 0.0f;
  for (long long base = rg.lo + threadIdx.x; base < rg.hi; base += kInvThreads * kInvVoxels) {
#pragma unroll
    for (int v = 0; v < kInvVoxels; ++v) {
      const long long g = base + v * kInvThreads;
      if (g < rg.hi) {
        int i, j, k;
        inv_coords(ix, g, i, j, k);
        const float wx = psi[g], wy = psi[g + n], wz = psi[g + 2 * n];
        float ex = res[g], ey = res[g + n], ez = res[g + 2 * n];
        const float nrm = scaled_norm(inv_sp, ex, ey, ez);
        float ux = -ex, uy = -ey, uz = -ez;
        if (nrm > limit) {
          const float f = limit / nrm;
          ux *= f;
          uy *= f;
          uz *= f;
        }
        float nx = wx + ux * eps, ny = wy + uy * eps, nz = wz + uz * eps;
        if (enforce && (i == 0 || j == 0 || k == 0 || i == d.nx - 1 || j == d.ny - 1 || k == d.nz - 1)) nx = ny = nz = 0.0f;
        psi[g] = nx;
        psi[g + n] = ny;
        psi[g + 2 * n] = nz;
        if (next) {
          inv_residual_at(field, d, l.v, i, j, k, nx, ny, nz, ex, ey, ez);
          res[g] = ex;
          res[g + n] = ey;
          res[g + 2 * n] = ez;
          const float r = scaled_norm(inv_sp, ex, ey, ez);
          m = nan_max(m, r);
          kahan_add(s, c, r);
        }
      }
    }
  }
  if (next) inv_write_partials(s, m, out, blocks, rg.chunk);
}

__global__ __launch_bounds__(kInvThreads) void k_inv_begin(const float* __restrict__ field, const float* __restrict__ psi,
                                                           float* __restrict__ res, Dims d, InvIndex ix, L9 l, float3 inv_sp,
                                                           InvertState* state, InvertReport* report, float* partials,
                                                           int max_iterations) {
  inv_begin_block(field, psi, res, d, ix, l, inv_sp, state, report, partials, max_iterations, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kInvThreads) void k_inv_iter(const float* __restrict__ field, float* __restrict__ psi,
                                                          float* __restrict__ res, Dims d, InvIndex ix, L9 l, float3 inv_sp,
                                                          InvertState* state, InvertReport* report, float* partials, int step,
                                                          int max_iterations, float max_tol, float mean_tol, int enforce,
                                                          uint32_t* mailbox, uint32_t tag) {
  inv_iter_block(field, psi, res, d, ix, l, inv_sp, state, report, partials, step, max_iterations, max_tol, mean_tol, enforce,
                 mailbox, mailbox, tag, blockIdx.x, gridDim.x);
}

// Several inversions on one grid: item q takes blocks [q * blocks, (q + 1) * blocks), the blocks
// of its own launch, and stores its stop hint in mailbox[kInvItemWords + q].
__global__ __launch_bounds__(kInvThreads) void k_inv_begin_items(InvItems it, Dims d, InvIndex ix, L9 l, float3 inv_sp,
                                                                 int blocks, int max_iterations) {
  const unsigned q = blockIdx.x / (unsigned)blocks;
  inv_begin_block(it.field[q], it.psi[q], it.res[q], d, ix, l, inv_sp, it.state[q], it.report[q], it.partials[q],
                  max_iterations, blockIdx.x - q * blocks, blocks);
}

__global__ __launch_bounds__(kInvThreads) void k_inv_iter_items(InvItems it, Dims d, InvIndex ix, L9 l, float3 inv_sp,
                                                                int blocks, int step, int max_iterations, float max_tol,
                                                                float mean_tol, int enforce, uint32_t* mailbox,
                                                                uint32_t tag) {
  const unsigned q = blockIdx.x / (unsigned)blocks;
  inv_iter_block(it.field[q], it.psi[q], it.res[q], d, ix, l, inv_sp, it.state[q], it.report[q], it.partials[q], step,
                 max_iterations, max_tol, mean_tol, enforce, mailbox, mailbox + kInvItemWords + q, tag,
                 blockIdx.x - q * blocks, blocks);
}

int inv_blocks(const Dims& d, int cap) {
  const long long b = (d.n() + kInvThreads * kInvVoxels - 1) / (kInvThreads * kInvVoxels);
  const long long c = cap < 1 ? 1 : cap;
  return (int)(b < 1 ? 1 : (b > c ? c : b));
}

void launch_inv_begin(const float* field, const float* psi, float* residual, Dims d, const float l[9], float3 inv_spacing,
                      InvertState* state, InvertReport* report, float* partials, int blocks, int max_iterations,
                      hipStream_t s) {
  L9 lv;
  for (int q = 0; q < 9; ++q) lv.v[q] = l[q];
  launch(k_inv_begin, dim3(blocks), dim3(kInvThreads), 0, s, field, psi, residual, d, inv_index(d), lv, inv_spacing, state,
         report, partials, max_iterations);
}

void launch_inv_begin_items(const InvItems& it, Dims d, const float l[9], float3 inv_spacing, int blocks, int max_iterations,
                            hipStream_t s) {
  if (it.count < 1 || it.count > kBatchFields) return;
  L9 lv;
  for (int q = 0; q < 9; ++q) lv.v[q] = l[q];
  launch(k_inv_begin_items, dim3(blocks * it.count), dim3(kInvThreads), 0, s, it, d, inv_index(d), lv, inv_spacing, blocks,
         max_iterations);
}

void launch_inv_iter_items(const InvItems& it, Dims d, const float l[9], float3 inv_spacing, int blocks, int step,
                           int max_iterations, double max_tol, double mean_tol, int enforce_boundary, uint32_t* mailbox,
                           uint32_t tag, hipStream_t s) {
  if (it.count < 1 || it.count > kBatchFields) return;
  L9 lv;
  for (int q = 0; q < 9; ++q) lv.v[q] = l[q];
  launch(k_inv_iter_items, dim3(blocks * it.count), dim3(kInvThreads), 0, s, it, d, inv_index(d), lv, inv_spacing, blocks,
         step, max_iterations, (float)max_tol, (float)mean_tol, enforce_boundary, mailbox, tag);
}

void launch_inv_iter(const float* field, float* psi, float* residual, Dims d, const float l[9], float3 inv_spacing,
                     InvertState* state, InvertReport* report, float* partials, int blocks, int step, int max_iterations,
                     double max_tol, double mean_tol, int enforce_boundary, uint32_t* mailbox, uint32_t tag, hipStream_t s) {
  L9 lv;
  for (int q = 0; q < 9; ++q) lv.v[q] = l[q];
  launch(k_inv_iter, dim3(blocks), dim3(kInvThreads), 0, s, field, psi, residual, d, inv_index(d), lv, inv_spacing, state,
         report, partials, step, max_iterations, (float)max_tol, (float)mean_tol, enforce_boundary, mailbox, tag);
}


}  // namespace fw::hip
