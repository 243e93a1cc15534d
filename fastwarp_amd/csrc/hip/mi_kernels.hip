// Mattes mutual information on gfx950: the sampled metric of the linear stages and the dense pair
// of a SyN MI channel (crates/fastwarp-reg/src/metric/mi.rs:364-676).
//
// Histograms: each cubic B-spline Parzen weight (f32) is converted exactly to 2^-32 fixed point and
// added with a 64-bit LDS atomic into a block-private histogram, which is then added into the
// global histogram with 64-bit atomics. Integer addition is associative, so the histogram is the
// same bits whatever the block count or scheduling. The host finishes it in f64.
// Parameter derivatives of the linear stage accumulate in FP64 per thread and reduce per block in
// a fixed tree; the host adds the block partials in block order.
#include "device_common.hpp"
#include "mi_kernels.hpp"

namespace fw::hip {

namespace {

// mi.rs:62-74 (ITK BSplineKernelFunction<3>)
__device__ __forceinline__ float bspline3(float u) {
  const float a = fabsf(u);
  if (a < 1.0f) {
    const float s = a * a;
    return (4.0f - 6.0f * s + 3.0f * s * a) * (1.0f / 6.0f);
  }
  if (a < 2.0f) {
    const float s = a * a;
    return (8.0f - 12.0f * a + 6.0f * s - s * a) * (1.0f / 6.0f);
  }
  return 0.0f;
}

// mi.rs:77-88 (ITK BSplineDerivativeKernelFunction<3>)
__device__ __forceinline__ float bspline3_d(float u) {
  if (u >= 0.0f && u < 1.0f) return -2.0f * u + 1.5f * u * u;
  if (u > -1.0f && u < 0.0f) return -2.0f * u - 1.5f * u * u;
  if (u >= 1.0f && u < 2.0f) return -2.0f + 2.0f * u - 0.5f * u * u;
  if (u > -2.0f && u <= -1.0f) return 2.0f + 2.0f * u + 0.5f * u * u;
  return 0.0f;
}

__device__ __forceinline__ unsigned long long to_fixed(float w) {
  return (unsigned long long)__float2ull_rn(w * 4294967296.0f);
}

__device__ __forceinline__ int clamp_bin(float term, int bins) {
  int b = (int)term;  // truncation toward zero, as the reference's `as i64`
  return b < 2 ? 2 : (b > bins - 3 ? bins - 3 : b);
}

// Central difference along physical axis `a` at a continuous index (interp.rs:102-121).
__device__ __forceinline__ float central_difference_at(const float* __restrict__ img, const Dims& d, const Probes& pr, int a,
                                                       float cx, float cy, float cz) {
  const float px = cx + pr.o[a][0], py = cy + pr.o[a][1], pz = cz + pr.o[a][2];
  const float mx = cx - pr.o[a][0], my = cy - pr.o[a][1], mz = cz - pr.o[a][2];
  if (!inside(d, px, py, pz) || !inside(d, mx, my, mz)) return 0.0f;
  return (tri(img, d, px, py, pz) - tri(img, d, mx, my, mz)) * pr.inv_spacing[a];
}

// The moving-image gradient of the linear stage's derivative at a continuous index: the three
// central differences of interp.rs:102-121. On a grid whose probes lie on the grid axes (a signed
// permutation of the axes) and where the 3 x 3 x 3 block from floor(c - 0.5) lies inside the
// image, every probe's trilinear stencil lies in that block. Its 27 values are loaded once. A
// linear interpolation at a point of the block is a product of per-axis weight triples (1 - w, w, 0)
// or (0, 1 - w, w), so the block is first contracted over the two axes the probes share with the
// centre, and each difference is the third axis's weight difference applied to that contraction
// (the same values as interpolating each probe, up to rounding order). Elsewhere the probes are
// interpolated one by one.
__device__ __forceinline__ void axis_triple(float c, int s, float w3[3]) {
  const float f = floorf(c);
  const bool first = (int)f == s;
  const float dd = c - f;
  const float w = dd > 0.0f ? dd : 0.0f;
  w3[0] = first ? 1.0f - w : 0.0f;
  w3[1] = first ? w : 1.0f - w;
  w3[2] = first ? 0.0f : w;
}

__device__ __forceinline__ void moving_gradient(const float* __restrict__ img, const Dims& d, const Probes& pr, float cx,
                                                float cy, float cz, float& g0, float& g1, float& g2) {
  if (pr.diagonal) {
    const int sx = (int)floorf(cx - 0.5f), sy = (int)floorf(cy - 0.5f), sz = (int)floorf(cz - 0.5f);
    if (sx >= 0 && sy >= 0 && sz >= 0 && sx + 2 <= d.nx - 1 && sy + 2 <= d.ny - 1 && sz + 2 <= d.nz - 1) {
      float v[3][3][3];
      const long long ys = d.nx, zs = (long long)d.nx * d.ny;
      const float* p = img + (long long)sx + ys * sy + zs * sz;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int i = 0; i < 3; ++i) v[k][j][i] = p[i + ys * j + zs * k];
      float wx[3], wy[3], wz[3];
      axis_triple(cx, sx, wx);
      axis_triple(cy, sy, wy);
      axis_triple(cz, sz, wz);
      // contractions over two axes at the centre's weights: X over (y, z), Y over (x, z), Z over (x, y)
      float X[3], Y[3], Z[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int t = 0; t < 3; ++t) {
            ax = fmaf(wy[r] * wz[t], v[t][r][q], ax);
            ay = fmaf(wx[r] * wz[t], v[t][q][r], ay);
            az = fmaf(wx[r] * wy[t], v[q][t][r], az);
          }
        X[q] = ax;
        Y[q] = ay;
        Z[q] = az;
      }
      const float c3[3] = {cx, cy, cz};
      const int s3[3] = {sx, sy, sz};
      float g[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // the probes of physical axis a move along one grid axis
        const int ax = pr.o[a][0] != 0.0f ? 0 : (pr.o[a][1] != 0.0f ? 1 : 2);
        const float h = pr.o[a][ax];
        float hi[3], lo[3];
        axis_triple(c3[ax] + h, s3[ax], hi);
        axis_triple(c3[ax] - h, s3[ax], lo);
        const float* line = ax == 0 ? X : (ax == 1 ? Y : Z);
        const float dv = (hi[0] - lo[0]) * line[0] + (hi[1] - lo[1]) * line[1] + (hi[2] - lo[2]) * line[2];
        g[a] = dv * pr.inv_spacing[a];
      }
      g0 = g[0];
      g1 = g[1];
      g2 = g[2];
      return;
    }
  }
  g0 = central_difference_at(img, d, pr, 0, cx, cy, cz);
  g1 = central_difference_at(img, d, pr, 1, cx, cy, cz);
  g2 = central_difference_at(img, d, pr, 2, cx, cy, cz);
}

// The samples of block b: one contiguous range per logical block (xcd_slab). The samples are in
// raster order of the fixed grid, so a block's moving-image gathers stay in one region and
// neighbouring regions run on the same XCD.
struct SampleRange {
  unsigned chunk;
  int lo, hi;
};
__device__ __forceinline__ SampleRange sample_range(int count) {
  const int len = ((count + (int)gridDim.x - 1) / (int)gridDim.x + 255) / 256 * 256;
  SampleRange r;
  r.chunk = xcd_slab(blockIdx.x, gridDim.x);
  const long long lo = (long long)r.chunk * len;
  r.lo = lo < count ? (int)lo : count;
  r.hi = lo + len < count ? (int)(lo + len) : count;
  return r;
}

struct SampleEval {
  bool ok;
  int fbin, mbin;
  float term, cx, cy, cz;
};

__device__ __forceinline__ SampleEval eval_sample(const MiSampledArgs& a, float4 s) {
  SampleEval e;
  e.ok = false;
  e.fbin = __float_as_int(s.w);
  e.cx = fmaf(a.a[0], s.x, fmaf(a.a[1], s.y, fmaf(a.a[2], s.z, a.b[0])));
  e.cy = fmaf(a.a[3], s.x, fmaf(a.a[4], s.y, fmaf(a.a[5], s.z, a.b[1])));
  e.cz = fmaf(a.a[6], s.x, fmaf(a.a[7], s.y, fmaf(a.a[8], s.z, a.b[2])));
  if (!inside(a.md, e.cx, e.cy, e.cz)) return e;
  const float mv = tri(a.moving, a.md, e.cx, e.cy, e.cz);
  if (!(mv >= a.moving_min && mv <= a.moving_max)) return e;  // also rejects NaN
  e.term = (mv - a.moving_min) * a.inv_bin_size + 2.0f;
  e.mbin = clamp_bin(e.term, a.bins);
  e.ok = true;
  return e;
}

}  // namespace

// One sample per thread while that gives at most 512 blocks (the small levels: their launches are
// one chain of dependent loads per thread, so the chain is kept short), several per thread beyond.
int mi_sampled_blocks(int count, int cap) {
  int b = (count + 255) / 256;
  return b < 1 ? 1 : (b > cap ? cap : b);
}

#ifndef FW_MI_WIDE_THREADS
#define FW_MI_WIDE_THREADS 1024
#endif

// Threads per block of the histogram passes: 256 while that gives each thread at most two samples,
// wider beyond. A wider block keeps the number of block histograms (and of slots to add) and puts
// more loads in flight per CU: at 512 blocks of 256 threads a CU holds two waves per SIMD, and the
// passes are bound by the latency of the sample and moving-image loads.
int mi_sampled_threads(int count) { return count > 2 * 512 * 256 ? FW_MI_WIDE_THREADS : 256; }

// The derivative kernel's partials are reduced on the device, so its grid may be wider.
int mi_deriv_blocks(int count, int cap) {
  int b = (count + 255) / 256;
  const int c = cap < kMiDerivMaxBlocks ? cap : kMiDerivMaxBlocks;
  return b < 1 ? 1 : (b > c ? c : b);
}

// The block's histogram (LDS, 64-bit atomics) is written to its own slot of `hist_slots`; a second
// kernel adds the slots, so the global histogram needs neither clearing nor global atomics.
__global__ __launch_bounds__(1024) void k_mi_sampled_hist(MiSampledArgs a) {
  extern __shared__ unsigned long long h[];
  const int bins = a.bins;
  const int cells = bins * bins + bins;
  for (int i = threadIdx.x; i < cells + 1; i += blockDim.x) h[i] = 0ull;
  __syncthreads();
  unsigned int mine = 0;
  const SampleRange rg = sample_range(a.count);
  for (int s = rg.lo + threadIdx.x; s < rg.hi; s += blockDim.x) {
    const SampleEval e = eval_sample(a, a.samples[s]);
    if (!e.ok) continue;
    ++mine;
    atomicAdd(&h[bins * bins + e.fbin], 1ull);
    unsigned long long* row = h + e.fbin * bins;
#pragma unroll
    for (int q = -1; q <= 2; ++q) {
      const int m = e.mbin + q;
      const float w = bspline3((float)m - e.term);
      if (w > 0.0f) atomicAdd(&row[m], to_fixed(w));
    }
  }
  if (mine) atomicAdd(&h[cells], (unsigned long long)mine);
  __syncthreads();
  unsigned long long* slot = a.hist_slots + (size_t)blockIdx.x * (cells + 1);
  for (int i = threadIdx.x; i < cells + 1; i += blockDim.x) slot[i] = h[i];
}

// K transforms per pass: each sample record is read once and binned under every transform, into
// the K histograms of the block's LDS (the samples, the anchor and the moving image are shared).
template <int K>
__global__ __launch_bounds__(1024) void k_mi_sampled_hist_batch(MiSampledArgs a, MiBatch batch) {
  extern __shared__ unsigned long long h[];
  const int bins = a.bins;
  const int cells = bins * bins + bins;
  const int C = cells + 1;
  for (int i = threadIdx.x; i < K * C; i += blockDim.x) h[i] = 0ull;
  __syncthreads();
  unsigned int mine[K];
#pragma unroll
  for (int k = 0; k < K; ++k) mine[k] = 0;
  const SampleRange rg = sample_range(a.count);
  for (int s = rg.lo + threadIdx.x; s < rg.hi; s += blockDim.x) {
    const float4 smp = a.samples[s];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      MiSampledArgs ak = a;
#pragma unroll
      for (int q = 0; q < 9; ++q) ak.a[q] = batch.a[k][q];
#pragma unroll
      for (int q = 0; q < 3; ++q) ak.b[q] = batch.b[k][q];
      const SampleEval e = eval_sample(ak, smp);
      if (!e.ok) continue;
      ++mine[k];
      unsigned long long* hk = h + k * C;
      atomicAdd(&hk[bins * bins + e.fbin], 1ull);
      unsigned long long* row = hk + e.fbin * bins;
#pragma unroll
      for (int q = -1; q <= 2; ++q) {
        const int m = e.mbin + q;
        const float w = bspline3((float)m - e.term);
        if (w > 0.0f) atomicAdd(&row[m], to_fixed(w));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (mine[k]) atomicAdd(&h[k * C + cells], (unsigned long long)mine[k]);
  __syncthreads();
  unsigned long long* slot = a.hist_slots + (size_t)blockIdx.x * K * C;
  for (int i = threadIdx.x; i < K * C; i += blockDim.x) slot[i] = h[i];
}

int mi_batch_capacity(int bins) {
  const size_t one = size_t(bins * bins + bins + 1) * sizeof(unsigned long long);
  const size_t fit = (size_t(64) << 10) / one;
  return int(fit < 1 ? 1 : (fit > size_t(kMiMaxBatch) ? kMiMaxBatch : fit));
}

void launch_mi_sampled_hist_batch(const MiSampledArgs& a, const MiBatch& batch, int count, int blocks, hipStream_t s) {
  const size_t lds = size_t(count) * size_t(a.bins * a.bins + a.bins + 1) * sizeof(unsigned long long);
  const dim3 t(a.threads > 0 ? a.threads : mi_sampled_threads(a.count));
  switch (count) {
    case 2: launch(k_mi_sampled_hist_batch<2>, dim3(blocks), t, (uint32_t)lds, s, a, batch); break;
    case 3: launch(k_mi_sampled_hist_batch<3>, dim3(blocks), t, (uint32_t)lds, s, a, batch); break;
    case 4: launch(k_mi_sampled_hist_batch<4>, dim3(blocks), t, (uint32_t)lds, s, a, batch); break;
    default: break;
  }
}

void launch_mi_sampled_hist(const MiSampledArgs& a, int blocks, hipStream_t s) {
  const size_t lds = size_t(a.bins * a.bins + a.bins + 1) * sizeof(unsigned long long);
  launch(k_mi_sampled_hist, dim3(blocks), dim3(a.threads > 0 ? a.threads : mi_sampled_threads(a.count)), lds, s, a);
}

// hist[c] = sum of the slots' cell c, for every cell and the valid count. A block of 16 waves
// takes 64 cells: lane = cell, wave w adds slots w, w + 16, ... in four independent chains (so the
// loads of one wave overlap), and the 16 wave sums are added in LDS.
__global__ __launch_bounds__(1024) void k_mi_sum_slots(const unsigned long long* __restrict__ slots, int nslots,
                                                       int cells, unsigned long long* __restrict__ hist, uint32_t* done,
                                                       uint32_t tag) {
  __shared__ unsigned long long part[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  unsigned long long t0 = 0ull, t1 = 0ull, t2 = 0ull, t3 = 0ull;
  if (c < cells) {
    int b = w;
    for (; b + 48 < nslots; b += 64) {
      t0 += slots[(size_t)b * cells + c];
      t1 += slots[(size_t)(b + 16) * cells + c];
      t2 += slots[(size_t)(b + 32) * cells + c];
      t3 += slots[(size_t)(b + 48) * cells + c];
    }
    for (; b < nslots; b += 16) t0 += slots[(size_t)b * cells + c];
  }
  part[w][lane] = (t0 + t1) + (t2 + t3);
  __syncthreads();
  if (w == 0 && c < cells) {
    unsigned long long t = 0ull;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += part[q][lane];
    hist[c] = t;
  }
  // the wave that wrote this block's cells releases them to the host with its completion word
  // (none when the sums stay on the device for a later kernel: the system-scope release would
  // write back this XCD's L2)
  if (done && threadIdx.x == 0) __hip_atomic_store(done + blockIdx.x, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

int mi_sum_slots_blocks(int cells) { return (cells + 63) / 64; }

void launch_mi_sum_slots(const unsigned long long* slots, int nslots, int cells, unsigned long long* hist, uint32_t* done,
                         uint32_t tag, hipStream_t s) {
  launch(k_mi_sum_slots, dim3(mi_sum_slots_blocks(cells)), dim3(1024), 0, s, slots, nslots, cells, hist, done, tag);
}

// The Mattes finish (mi.rs:702-769) of histogram `blockIdx.x` on the device: hist holds the
// histograms one after another, each bins*bins joint cells (2^-32 fixed point), the bins fixed
// marginal counts and the valid count. The joint total is the exact integer sum of the cells,
// rounded once; the cells and marginals are scaled by its inverse; the moving marginal adds each
// column in row order (one thread per column, beside the fixed marginal's logs, one per row); the
// value adds jv (log(jv / p_M) - log(p_F)) over the cells whose three
// terms exceed the f64 epsilon (per-thread sums, then the waves' sums in order). ratio (block 0,
// if given) receives n_factor log(jv / p_M) per cell in f32, zero elsewhere, as the derivative
// pass reads it. out[b] gets the value, the valid count and a status (1: no valid sample, 2: the
// joint cells sum to zero; the ratio is then all zero); then, if done is given, tag goes to
// done[b] with a system-scope release.
// The finish of one histogram h (global or LDS) of at most MAXB bins by the whole block (every
// thread calls it).
template <int MAXB>
__device__ void finish_block(const unsigned long long* h, int bins, double moving_bin_size, MiFinishOut* out, float* rb,
                             uint32_t* done, uint32_t tag) {
  __shared__ double jn[MAXB * MAXB];
  __shared__ double mm[MAXB], fm[MAXB], lf[MAXB];
  __shared__ unsigned long long wsum_u[16];
  __shared__ double wsum_d[16];
  __shared__ int status_s;
  const double eps = 2.220446049250313e-16;
  const int cells = bins * bins;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = (int)(blockDim.x + 63) >> 6;
  __syncthreads();  // the scratch of an earlier call in this block is free
  const unsigned long long valid = h[cells + bins];
  // the exact total of the joint cells
  unsigned long long t = 0ull;
  for (int c = tid; c < cells; c += blockDim.x) t += h[c];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  if (lane == 0) wsum_u[w] = t;
  __syncthreads();
  if (tid == 0) {
    unsigned long long all = 0ull;
    for (int q = 0; q < nw; ++q) all += wsum_u[q];
    const double sum = (double)all / kMiFixedPoint;
    status_s = valid == 0ull ? 1 : (sum < eps ? 2 : 0);
    wsum_d[0] = status_s == 0 ? 1.0 / sum : 0.0;
  }
  __syncthreads();
  const int status = status_s;
  const double norm = wsum_d[0];
  if (status != 0) {
    if (rb)
      for (int c = tid; c < cells; c += blockDim.x) rb[c] = 0.0f;
    if (tid == 0) {
      out->value = 0.0;
      out->valid = valid;
      out->status = status;
      if (done) __hip_atomic_store(done, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  for (int c = tid; c < cells; c += blockDim.x) jn[c] = ((double)h[c] / kMiFixedPoint) * norm;
  if (tid < bins) fm[tid] = (double)h[cells + tid] * norm;
  __syncthreads();
  // the fixed marginal's logs on the first wave, the moving marginal on the second
  if (tid < bins) {
    lf[tid] = fm[tid] > eps ? log(fm[tid]) : 0.0;
  } else if (tid >= 64 && tid < 64 + bins) {
    const int m = tid - 64;
    double s = 0.0;
#pragma unroll 8
    for (int r = 0; r < bins; ++r) s += jn[r * bins + m];
    mm[m] = s;
  }
  __syncthreads();
  const double n_factor = -1.0 / (moving_bin_size * (double)valid);
  double total = 0.0;
  for (int c = tid; c < cells; c += blockDim.x) {
    const int f = c / bins, m = c - f * bins;
    const double jv = jn[c];
    float r = 0.0f;
    if (fm[f] > eps && mm[m] > eps && jv > eps) {
      const double pr = log(jv / mm[m]);
      total += jv * (pr - lf[f]);
      r = (float)(n_factor * pr);
    }
    if (rb) rb[c] = r;
  }
  for (int off = 32; off > 0; off >>= 1) total += __shfl_down(total, off, 64);
  __syncthreads();
  if (lane == 0) wsum_d[w] = total;
  __syncthreads();
  if (tid == 0) {
    double all = 0.0;
    for (int q = 0; q < nw; ++q) all += wsum_d[q];
    out->value = -all;
    out->valid = valid;
    out->status = 0;
    if (done) __hip_atomic_store(done, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ __launch_bounds__(1024) void k_mi_finish(const unsigned long long* __restrict__ hist, int bins,
                                                    double moving_bin_size, MiFinishOut* out, float* __restrict__ ratio,
                                                    uint32_t* done, uint32_t tag) {
  const int b = blockIdx.x;
  finish_block<kMiMaxBins>(hist + (size_t)b * (bins * bins + bins + 1), bins, moving_bin_size, out + b,
                           b == 0 ? ratio : nullptr, done ? done + b : nullptr, tag);
}

void launch_mi_finish(const unsigned long long* hist, int count, int bins, double moving_bin_size, MiFinishOut* out,
                      float* ratio, uint32_t* done, uint32_t tag, hipStream_t s) {
  launch(k_mi_finish, dim3(count), dim3(1024), 0, s, hist, bins, moving_bin_size, out, ratio, done, tag);
}

// The slot sum and the finish in one launch, for passes of few slots and at most kMiFusedBins
// bins: block k adds histogram k's cells over the slots (slot q holds `per_slot` entries, the
// histograms one after another) into LDS, as k_mi_sum_slots does (integer sums, so the same
// cells), and finishes them.
__global__ __launch_bounds__(1024) void k_mi_sum_finish(const unsigned long long* __restrict__ slots, int nslots,
                                                        int per_slot, int bins, double moving_bin_size, MiFinishOut* out,
                                                        float* __restrict__ ratio, uint32_t* done, uint32_t tag) {
  __shared__ unsigned long long hs[kMiFusedBins * kMiFusedBins + kMiFusedBins + 1];
  const int C = bins * bins + bins + 1;
  const int b = blockIdx.x;
  const unsigned long long* src = slots + (size_t)b * C;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    unsigned long long t0 = 0ull, t1 = 0ull, t2 = 0ull, t3 = 0ull;
    int q = 0;
    for (; q + 3 < nslots; q += 4) {
      t0 += src[(size_t)q * per_slot + c];
      t1 += src[(size_t)(q + 1) * per_slot + c];
      t2 += src[(size_t)(q + 2) * per_slot + c];
      t3 += src[(size_t)(q + 3) * per_slot + c];
    }
    for (; q < nslots; ++q) t0 += src[(size_t)q * per_slot + c];
    hs[c] = (t0 + t1) + (t2 + t3);
  }
  finish_block<kMiFusedBins>(hs, bins, moving_bin_size, out + b, b == 0 ? ratio : nullptr, done ? done + b : nullptr, tag);
}

void launch_mi_sum_finish(const unsigned long long* slots, int nslots, int count, int bins, double moving_bin_size,
                          MiFinishOut* out, float* ratio, uint32_t* done, uint32_t tag, hipStream_t s) {
  const int per_slot = count * (bins * bins + bins + 1);
  launch(k_mi_sum_finish, dim3(count), dim3(1024), 0, s, slots, nslots, per_slot, bins, moving_bin_size, out, ratio, done,
         tag);
}

// Per sample: sum_k B3'(k - term) * ratio[f, k] * (grad M . dT/dtheta_p) for every parameter
// (mi.rs:405-433 with the ratio weights of :752-757), accumulated in FP64.
__global__ __launch_bounds__(256) void k_mi_sampled_deriv(MiSampledArgs a) {
  __shared__ float ratio[kMiMaxBins * kMiMaxBins];
  __shared__ double red[4 * 12];
  const int bins = a.bins;
  for (int i = threadIdx.x; i < bins * bins; i += blockDim.x) ratio[i] = a.ratio[i];
  __syncthreads();
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  const SampleRange rg = sample_range(a.count);
  for (int s = rg.lo + threadIdx.x; s < rg.hi; s += blockDim.x) {
    const float4 smp = a.samples[s];
    const SampleEval e = eval_sample(a, smp);
    if (!e.ok) continue;
    float w = 0.0f;
#pragma unroll
    for (int k = -1; k <= 2; ++k) {
      const int m = e.mbin + k;
      w += bspline3_d((float)m - e.term) * ratio[e.fbin * bins + m];
    }
    if (w == 0.0f) continue;
    float g0, g1, g2;
    moving_gradient(a.moving, a.md, a.probes, e.cx, e.cy, e.cz, g0, g1, g2);
    if (!(isfinite(g0) && isfinite(g1) && isfinite(g2))) g0 = g1 = g2 = 0.0f;
    // g' = outer^T grad: the Jacobian columns are left-multiplied by the outer matrix
    const float q0 = a.outer[0] * g0 + a.outer[3] * g1 + a.outer[6] * g2;
    const float q1 = a.outer[1] * g0 + a.outer[4] * g1 + a.outer[7] * g2;
    const float q2 = a.outer[2] * g0 + a.outer[5] * g1 + a.outer[8] * g2;
    const float px = smp.x + a.rel_center[0], py = smp.y + a.rel_center[1], pz = smp.z + a.rel_center[2];
    if (a.kind == 1) {
      // affine (param.rs:287-297): column r*3+c is e_r (p_c - c_c), column 9+r is e_r
      const float q[3] = {q0, q1, q2};
      const float p[3] = {px, py, pz};
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r * 3 + c] += (double)(w * q[r] * p[c]);
        acc[9 + r] += (double)(w * q[r]);
      }
    } else {
      // Euler3D, matrix Rz Rx Ry (param.rs:254-281)
      const float cx = a.trig[0], sx = a.trig[1], cy = a.trig[2], sy = a.trig[3], cz = a.trig[4], sz = a.trig[5];
      const float j00 = (-sz * cx * sy) * px + (sz * sx) * py + (sz * cx * cy) * pz;
      const float j01 = (cz * cx * sy) * px + (-cz * sx) * py + (-cz * cx * cy) * pz;
      const float j02 = (sx * sy) * px + cx * py + (-sx * cy) * pz;
      const float j10 = (-cz * sy - sz * sx * cy) * px + (cz * cy - sz * sx * sy) * pz;
      const float j11 = (-sz * sy + cz * sx * cy) * px + (sz * cy + cz * sx * sy) * pz;
      const float j12 = (-cx * cy) * px + (-cx * sy) * pz;
      const float j20 = (-sz * cy - cz * sx * sy) * px + (-cz * cx) * py + (-sz * sy + cz * sx * cy) * pz;
      const float j21 = (cz * cy - sz * sx * sy) * px + (-sz * cx) * py + (cz * sy + sz * sx * cy) * pz;
      acc[0] += (double)(w * (q0 * j00 + q1 * j01 + q2 * j02));
      acc[1] += (double)(w * (q0 * j10 + q1 * j11 + q2 * j12));
      acc[2] += (double)(w * (q0 * j20 + q1 * j21));
      acc[3] += (double)(w * q0);
      acc[4] += (double)(w * q1);
      acc[5] += (double)(w * q2);
    }
  }
  // fixed tree: a butterfly within each wave, then the four wave sums in order
  const int np = a.kind == 1 ? 12 : 6;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    if (k >= np) break;
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[w * 12 + k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < np) {
    const int k = threadIdx.x;
    a.partial[rg.chunk * 12 + k] = (red[k] + red[12 + k]) + (red[24 + k] + red[36 + k]);
  }
}

void launch_mi_sampled_deriv(const MiSampledArgs& a, int blocks, hipStream_t s) {
  launch(k_mi_sampled_deriv, dim3(blocks), dim3(256), 0, s, a);
}

// out[k] = sum over the blocks of partial[b * 12 + k], k < np, in a fixed order: 16 waves take
// consecutive slices of the blocks, each lane adds its blocks' rows (all loads in flight at
// once), a butterfly adds each wave's lanes, and the 16 wave sums are added in order.
__global__ __launch_bounds__(1024) void k_mi_deriv_finish(const double* __restrict__ partial, int blocks, int np,
                                                          double* __restrict__ out, uint32_t* done, uint32_t tag,
                                                          const MiFinishOut* finish, MiFinishOut* finish_out) {
  __shared__ double wsum[16][12];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int per = (blocks + 15) / 16;
  const int b0 = w * per, b1 = min(blocks, b0 + per);
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  for (int b = b0 + lane; b < b1; b += 64) {
    const double2* row = reinterpret_cast<const double2*>(partial + (size_t)b * 12);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double2 v = row[q];
      acc[2 * q] += v.x;
      acc[2 * q + 1] += v.y;
    }
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) wsum[w][k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < np) {
    double t = 0.0;
    for (int q = 0; q < 16; ++q) t += wsum[q][threadIdx.x];
    out[threadIdx.x] = t;
  }
  // wave 0 wrote the sums (and copies the value pass's finish beside them); it releases them to
  // the host with the completion word
  if (threadIdx.x == 0) {
    if (finish) {
      finish_out->value = finish->value;
      finish_out->valid = finish->valid;
      finish_out->status = finish->status;
    }
    __hip_atomic_store(done, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

void launch_mi_deriv_finish(const double* partial, int blocks, int np, double* out, uint32_t* done, uint32_t tag,
                            hipStream_t s, const MiFinishOut* finish, MiFinishOut* finish_out) {
  launch(k_mi_deriv_finish, dim3(1), dim3(1024), 0, s, partial, blocks, np, out, done, tag, finish, finish_out);
}

// ------------------------------------------------------------------------------ dense pair ---

int mi_dense_blocks(long long n) {
  long long b = (n + 4095) / 4096;
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

__device__ __forceinline__ bool dense_inside(const MiDenseArgs& a, long long g) {
  return (!a.fmask || a.fmask[g] != 0.0f) && (!a.mmask || a.mmask[g] != 0.0f);
}

// mi.rs:171-184 / 290-304: min and max of each image over the voxels inside its own mask.
__global__ __launch_bounds__(256) void k_mi_dense_range(MiDenseArgs a) {
  const long long n = a.d.n();
  float fmin = INFINITY, fmax = -INFINITY, mmin = INFINITY, mmax = -INFINITY;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
    const float f = a.f[g], m = a.m[g];
    if ((!a.fmask || a.fmask[g] != 0.0f) && isfinite(f)) {
      fmin = fminf(fmin, f);
      fmax = fmaxf(fmax, f);
    }
    if ((!a.mmask || a.mmask[g] != 0.0f) && isfinite(m)) {
      mmin = fminf(mmin, m);
      mmax = fmaxf(mmax, m);
    }
  }
  __shared__ float r[4][256];
  r[0][threadIdx.x] = fmin;
  r[1][threadIdx.x] = fmax;
  r[2][threadIdx.x] = mmin;
  r[3][threadIdx.x] = mmax;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r[0][threadIdx.x] = fminf(r[0][threadIdx.x], r[0][threadIdx.x + st]);
      r[1][threadIdx.x] = fmaxf(r[1][threadIdx.x], r[1][threadIdx.x + st]);
      r[2][threadIdx.x] = fminf(r[2][threadIdx.x], r[2][threadIdx.x + st]);
      r[3][threadIdx.x] = fmaxf(r[3][threadIdx.x], r[3][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) a.range_partial[blockIdx.x * 4 + threadIdx.x] = r[threadIdx.x][0];
}

void launch_mi_dense_range(const MiDenseArgs& a, int blocks, hipStream_t s) {
  launch(k_mi_dense_range, dim3(blocks), dim3(256), 0, s, a);
}

__device__ __forceinline__ int threshold_bin(const float* thr, int bins, float v) {
  int b = 2;
  for (int q = 0; q < bins - 5; ++q) b += (v >= thr[q]) ? 1 : 0;  // thresholds of bins 3 .. bins-3
  return b;
}

// Both directions' joint histograms and fixed marginals in one pass.
__global__ __launch_bounds__(256) void k_mi_dense_hist(MiDenseArgs a) {
  extern __shared__ unsigned long long hs[];
  const int bins = a.bins;
  const int stride = bins * bins + bins;
  for (int i = threadIdx.x; i < 2 * stride; i += blockDim.x) hs[i] = 0ull;
  __shared__ unsigned int valid[2];
  if (threadIdx.x < 2) valid[threadIdx.x] = 0u;
  __syncthreads();
  const long long n = a.d.n();
  unsigned int mine[2] = {0u, 0u};
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
    if (!dense_inside(a, g)) continue;
    const float f = a.f[g], m = a.m[g];
    if (!isfinite(f) || !isfinite(m)) continue;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      const float fixed = dir == 0 ? f : m;
      const float moving = dir == 0 ? m : f;
      if (!(moving >= a.moving_min[dir] && moving <= a.moving_max[dir])) continue;
      const float term = (moving - a.moving_min[dir]) * a.inv_bin_size[dir] + 2.0f;
      const int mbin = clamp_bin(term, bins);
      const int fbin = threshold_bin(a.fixed_thr[dir], bins, fixed);
      unsigned long long* h = hs + dir * stride;
      ++mine[dir];
      atomicAdd(&h[bins * bins + fbin], 1ull);
#pragma unroll
      for (int k = -1; k <= 2; ++k) {
        const int mb = mbin + k;
        const float w = bspline3((float)mb - term);
        if (w > 0.0f) atomicAdd(&h[fbin * bins + mb], to_fixed(w));
      }
    }
  }
  if (mine[0]) atomicAdd(&valid[0], mine[0]);
  if (mine[1]) atomicAdd(&valid[1], mine[1]);
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * stride; i += blockDim.x) {
    const int dir = i / stride, cell = i % stride;
    if (hs[i]) atomicAdd(&a.hist[dir * (stride + 1) + cell], hs[i]);
  }
  if (threadIdx.x < 2 && valid[threadIdx.x])
    atomicAdd(&a.hist[threadIdx.x * (stride + 1) + stride], (unsigned long long)valid[threadIdx.x]);
}

void launch_mi_dense_hist(const MiDenseArgs& a, int blocks, hipStream_t s) {
  const size_t shmem = size_t(2) * (a.bins * a.bins + a.bins) * sizeof(unsigned long long);
  launch(k_mi_dense_hist, dim3(blocks), dim3(256), shmem, s, a);
}

// Per voxel sum_k B3'(m_k - term) ratio[f, m_k] * grad(moving) for both directions (mi.rs:650-666).
__global__ __launch_bounds__(256) void k_mi_dense_deriv(MiDenseArgs a) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.z;
  if (i >= a.d.nx || j >= a.d.ny) return;
  const long long g = lin(a.d, i, j, k), n = a.d.n();
  float out[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  const float f = a.f[g], m = a.m[g];
  if (dense_inside(a, g) && isfinite(f) && isfinite(m)) {
    const int bins = a.bins;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      const float fixed = dir == 0 ? f : m;
      const float moving = dir == 0 ? m : f;
      if (!(moving >= a.moving_min[dir] && moving <= a.moving_max[dir])) continue;
      const float term = (moving - a.moving_min[dir]) * a.inv_bin_size[dir] + 2.0f;
      const int mbin = clamp_bin(term, bins);
      const int fbin = threshold_bin(a.fixed_thr[dir], bins, fixed);
      const float* ratio = a.ratio + dir * bins * bins + fbin * bins;
      float s = 0.0f;
#pragma unroll
      for (int q = -1; q <= 2; ++q) s += bspline3_d((float)(mbin + q) - term) * ratio[mbin + q];
      const float* img = dir == 0 ? a.m : a.f;
      float g0 = central_difference(img, a.d, a.probes, 0, i, j, k);
      float g1 = central_difference(img, a.d, a.probes, 1, i, j, k);
      float g2 = central_difference(img, a.d, a.probes, 2, i, j, k);
      if (!(isfinite(g0) && isfinite(g1) && isfinite(g2))) g0 = g1 = g2 = 0.0f;
      out[dir][0] = s * g0;
      out[dir][1] = s * g1;
      out[dir][2] = s * g2;
    }
  }
  a.d0[g] = out[0][0];
  a.d0[g + n] = out[0][1];
  a.d0[g + 2 * n] = out[0][2];
  a.d1[g] = out[1][0];
  a.d1[g + n] = out[1][1];
  a.d1[g + 2 * n] = out[1][2];
}

void launch_mi_dense_deriv(const MiDenseArgs& a, hipStream_t s) {
  const dim3 grid((a.d.nx + 63) / 64, (a.d.ny + 3) / 4, a.d.nz);
  launch(k_mi_dense_deriv, grid, dim3(256), 0, s, a);
}

}  // namespace fw::hip
