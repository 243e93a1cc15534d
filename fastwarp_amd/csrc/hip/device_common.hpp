// Device helpers shared by the kernel files: ITK interpolation and buffer rules, central
// differences, wavefront and block reductions, window counts.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace fw::hip {

namespace {

constexpr float kEps64 = 2.220446049250313e-16f;  // f64 machine epsilon, the reference's guard

// The workgroups of a dispatch go to the 8 XCDs of the MI355X in turn (workgroup b to XCD b % 8),
// and every XCD has an L2 of its own. xcd_slab renumbers them so that XCD x owns one contiguous
// range of logical ids: tiles that are neighbours in memory then run on one XCD, and a stencil or
// gather that reaches into the neighbouring tile finds the lines in that XCD's L2 instead of
// fetching them from HBM once more. A bijection on [0, n) whatever the placement really is.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_slab(unsigned b, unsigned n) {
  const unsigned per = n / kXcds, extra = n % kXcds;
  const unsigned x = b % kXcds, k = b / kXcds;
  return x * per + (x < extra ? x : extra) + k;
}

// The logical (x, y, z) block of a 3D grid under xcd_slab (x fastest, as the dispatch order).
__device__ __forceinline__ uint3 xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned l = xcd_slab(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  return make_uint3(l % gx, (l / gx) % gy, l / (gx * gy));
}

__device__ __forceinline__ long long lin(const Dims& d, int i, int j, int k) {
  return (long long)i + (long long)d.nx * ((long long)j + (long long)d.ny * k);
}

// ITK IsInsideBuffer on a continuous index: every coordinate in [-0.5, n - 0.5).
__device__ __forceinline__ bool inside(const Dims& d, float cx, float cy, float cz) {
  return cx >= -0.5f && cx < (float)d.nx - 0.5f && cy >= -0.5f && cy < (float)d.ny - 0.5f && cz >= -0.5f &&
         cz < (float)d.nz - 0.5f;
}

// ITK LinearInterpolateImageFunction::EvaluateOptimized (crates/fastwarp-reg/src/interp.rs:43-87):
// base clamped to the buffer, no interpolation along an axis whose fraction is not positive or
// whose upper neighbour is past the last voxel.
__device__ __forceinline__ void axis_weight(float c, int n, int& b, float& w) {
  const float f = floorf(c);
  int bb = f < 0.0f ? 0 : (int)f;
  bb = bb > n - 1 ? n - 1 : bb;
  const float dd = c - (float)bb;
  b = bb;
  w = (dd <= 0.0f || bb + 1 > n - 1) ? 0.0f : dd;
}

__device__ __forceinline__ float tri(const float* __restrict__ img, const Dims& d, float cx, float cy, float cz) {
  int bx, by, bz;
  float wx, wy, wz;
  axis_weight(cx, d.nx, bx, wx);
  axis_weight(cy, d.ny, by, wy);
  axis_weight(cz, d.nz, bz, wz);
  // all eight loads stay inside the buffer (a missing upper neighbour repeats the base, and its
  // weight is zero), so they are issued together and the blends select instead of branching
  const long long sx = bx + 1 < d.nx ? 1 : 0;
  const long long sy = by + 1 < d.ny ? (long long)d.nx : 0;
  const long long sz = bz + 1 < d.nz ? (long long)d.nx * d.ny : 0;
  const float* p = img + lin(d, bx, by, bz);
  const float v0 = p[0], v1 = p[sx], v2 = p[sy], v3 = p[sy + sx];
  const float v4 = p[sz], v5 = p[sz + sx], v6 = p[sz + sy], v7 = p[sz + sy + sx];
  const float r00 = wx > 0.0f ? v0 + (v1 - v0) * wx : v0;
  const float r01 = wx > 0.0f ? v2 + (v3 - v2) * wx : v2;
  const float r10 = wx > 0.0f ? v4 + (v5 - v4) * wx : v4;
  const float r11 = wx > 0.0f ? v6 + (v7 - v6) * wx : v6;
  const float p0 = wy > 0.0f ? r00 + (r01 - r00) * wy : r00;
  const float p1 = wy > 0.0f ? r10 + (r11 - r10) * wy : r10;
  return wz > 0.0f ? p0 + (p1 - p0) * wz : p0;
}

// All three planes of a field at one continuous index (the same rule as tri()): the eight corner
// offsets and weights are computed once and all 24 loads are issued before any blend.
__device__ __forceinline__ void tri3(const float* __restrict__ u, const Dims& d, float cx, float cy, float cz,
                                     float& ox, float& oy, float& oz) {
  int bx, by, bz;
  float wx, wy, wz;
  axis_weight(cx, d.nx, bx, wx);
  axis_weight(cy, d.ny, by, wy);
  axis_weight(cz, d.nz, bz, wz);
  const long long n = d.n();
  const long long sx = bx + 1 < d.nx ? 1 : 0;
  const long long sy = by + 1 < d.ny ? (long long)d.nx : 0;
  const long long sz = bz + 1 < d.nz ? (long long)d.nx * d.ny : 0;
  const long long g = lin(d, bx, by, bz);
  float v[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float* p = u + c * n + g;
    v[c][0] = p[0];
    v[c][1] = p[sx];
    v[c][2] = p[sy];
    v[c][3] = p[sy + sx];
    v[c][4] = p[sz];
    v[c][5] = p[sz + sx];
    v[c][6] = p[sz + sy];
    v[c][7] = p[sz + sy + sx];
  }
  float r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float r00 = wx > 0.0f ? v[c][0] + (v[c][1] - v[c][0]) * wx : v[c][0];
    const float r01 = wx > 0.0f ? v[c][2] + (v[c][3] - v[c][2]) * wx : v[c][2];
    const float r10 = wx > 0.0f ? v[c][4] + (v[c][5] - v[c][4]) * wx : v[c][4];
    const float r11 = wx > 0.0f ? v[c][6] + (v[c][7] - v[c][6]) * wx : v[c][6];
    const float p0 = wy > 0.0f ? r00 + (r01 - r00) * wy : r00;
    const float p1 = wy > 0.0f ? r10 + (r11 - r10) * wy : r10;
    r[c] = wz > 0.0f ? p0 + (p1 - p0) * wz : p0;
  }
  ox = r[0];
  oy = r[1];
  oz = r[2];
}

// Nearest voxel with rounding half up (crates/fastwarp-reg/src/mask.rs:99-111); -1 outside.
__device__ __forceinline__ long long nearest_index(const Dims& d, float cx, float cy, float cz) {
  const float rx = floorf(cx + 0.5f), ry = floorf(cy + 0.5f), rz = floorf(cz + 0.5f);
  if (!(rx >= 0.0f && rx < (float)d.nx && ry >= 0.0f && ry < (float)d.ny && rz >= 0.0f && rz < (float)d.nz)) return -1;
  return lin(d, (int)rx, (int)ry, (int)rz);
}

// One component of ITK's CentralDifferenceImageFunction at voxel (i, j, k) (interp.rs:102-121):
// probes at +-half a spacing along physical axis `a`, zero when either probe leaves the buffer.
__device__ __forceinline__ float central_difference(const float* __restrict__ img, const Dims& d, const Probes& pr,
                                                    int a, int i, int j, int k) {
  if (pr.diagonal) {
    // The probes lie on grid axis `ax` at -+0.5: interpolation reduces to its two neighbours.
    int ax = 0;
    float sgn = pr.o[a][0];
    if (pr.o[a][1] != 0.0f) {
      ax = 1;
      sgn = pr.o[a][1];
    } else if (pr.o[a][2] != 0.0f) {
      ax = 2;
      sgn = pr.o[a][2];
    }
    const int pos = ax == 0 ? i : ax == 1 ? j : k;
    const int n = ax == 0 ? d.nx : ax == 1 ? d.ny : d.nz;
    if (pos >= n - 1) return 0.0f;  // the upper probe is at n - 0.5, outside the buffer
    const long long stride = ax == 0 ? 1 : ax == 1 ? (long long)d.nx : (long long)d.nx * d.ny;
    const float* p = img + lin(d, i, j, k);
    const float c0 = p[0];
    const float up = c0 + (p[stride] - c0) * 0.5f;
    float down = c0;
    if (pos >= 1) {
      const float m1 = p[-stride];
      down = m1 + (c0 - m1) * 0.5f;
    }
    const float diff = sgn > 0.0f ? up - down : down - up;
    return diff * pr.inv_spacing[a];
  }
  const float ci = (float)i, cj = (float)j, ck = (float)k;
  const float px = ci + pr.o[a][0], py = cj + pr.o[a][1], pz = ck + pr.o[a][2];
  const float mx = ci - pr.o[a][0], my = cj - pr.o[a][1], mz = ck - pr.o[a][2];
  if (!inside(d, px, py, pz) || !inside(d, mx, my, mz)) return 0.0f;
  return (tri(img, d, px, py, pz) - tri(img, d, mx, my, mz)) * pr.inv_spacing[a];
}


__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ float nan_max(float a, float b) { return (b > a || b != b) ? b : a; }

__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_down(v, off, 64));
  return v;
}

// Block sum (blockDim.x a multiple of 64, at most 1024) in a fixed tree; result in thread 0.
__device__ float block_sum(float v) {
  __shared__ float red[16];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
  return t;
}

__device__ float block_max(float v) {
  __shared__ float red[16];
  v = wave_max(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t = nan_max(t, red[q]);
  return t;
}

__device__ __forceinline__ int window_count(int pos, int n, int r) {
  const int lo = pos > r ? pos - r : 0;
  const int hi = pos + r < n - 1 ? pos + r : n - 1;
  return hi - lo + 1;
}


}  // namespace

}  // namespace fw::hip
