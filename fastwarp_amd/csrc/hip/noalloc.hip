// fw_cc_pair_noalloc: the CC pair of fw_cc_pair on caller-owned device buffers, with no device or
// host allocation. Its kernels keep their block partials and the two image bounds in `scratch`, a
// static device array of fixed size (each GPU the library runs on holds its own copy until the
// process ends). Host code never passes the array's name to a kernel: the name is a host-side
// stand-in, and each device's address of its copy comes from hipGetSymbolAddress, taken once per
// device with that device current and kept in `g_scratch`.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU (on different
// streams, or from different threads) race on it. Every kernel of a call writes each scratch word
// it reads before reading it, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <mutex>
#include <string>

#include "../../../include/fastwarp_hip.h"
#include "../host/grid.hpp"
#include "device_common.hpp"
#include "geometry.hpp"
#include "backend.hpp"
#include "kernels.hpp"

namespace fw {
void set_call_error(const std::string& message);
}

namespace fw::hip {
namespace {

// The largest CC grid the scratch holds, in blocks of the CC kernel (64 x 8 x tile_z voxels): the
// 1 mm grid (182 x 218 x 182) takes 1008 blocks.
constexpr int kNoAllocMaxBlocks = 16384;
// Blocks of the bound pass.
constexpr int kBoundBlocks = 512;
// Layout in 4-byte words: CC value partials (f32), CC counts (i32), bound partials (2 per block,
// f32), then the two bounds.
constexpr size_t kOffCc = 0;
constexpr size_t kOffCount = kOffCc + kNoAllocMaxBlocks;
constexpr size_t kOffBoundPart = kOffCount + kNoAllocMaxBlocks;
constexpr size_t kOffBounds = kOffBoundPart + 2 * kBoundBlocks;
constexpr size_t kScratchWords = kOffBounds + 2;

// 137,224 bytes per device.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

// Per block: the largest finite |value| of each image (the bound fw_image_upload takes on the host).
__global__ __launch_bounds__(256) void k_bound_partials(const float* __restrict__ f, const float* __restrict__ m, long long n,
                                                        float* partial) {
  float a = 0.0f, b = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float x = f[i], y = m[i];
    if (isfinite(x)) a = fmaxf(a, fabsf(x));
    if (isfinite(y)) b = fmaxf(b, fabsf(y));
  }
  const float ra = block_max(a);
  const float rb = block_max(b);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = ra;
    partial[2 * blockIdx.x + 1] = rb;
  }
}

__global__ __launch_bounds__(256) void k_bound_final(const float* __restrict__ partial, int blocks, float* bounds) {
  float a = 0.0f, b = 0.0f;
  for (int i = threadIdx.x; i < blocks; i += blockDim.x) {
    a = fmaxf(a, partial[2 * i]);
    b = fmaxf(b, partial[2 * i + 1]);
  }
  const float ra = block_max(a);
  const float rb = block_max(b);
  if (threadIdx.x == 0) {
    bounds[0] = ra;
    bounds[1] = rb;
  }
}

// values[0] = values[1] = -(sum of the block partials) / (sum of the counts), the block partials
// added in block order in f64 as fw_cc_pair adds them on the host.
__global__ __launch_bounds__(64) void k_cc_value_final(const float* __restrict__ partial_cc, const int* __restrict__ count,
                                                       int blocks, double* values) {
  if (threadIdx.x != 0) return;
  double total = 0.0;
  long long c = 0;
  for (int b = 0; b < blocks; ++b) {
    total += (double)partial_cc[b];
    c += count[b];
  }
  const double v = c > 0 ? -total / (double)c : __builtin_nan("");
  values[0] = v;
  values[1] = v;
}

// Each device's address of its `scratch`, taken once with that device current.
std::mutex g_scratch_mutex;
std::array<uint32_t*, 64> g_scratch{};

uint32_t* scratch_of_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("fw_cc_pair_noalloc: hipGetDevice failed");
  if (dev < 0 || dev >= int(g_scratch.size())) throw std::runtime_error("fw_cc_pair_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("fw_cc_pair_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<uint32_t*>(p);
  }
  return g_scratch[dev];
}

}  // namespace
}  // namespace fw::hip

extern "C" int fw_cc_pair_noalloc(const fw_grid* grid, const float* fixed, const float* moving, uint32_t radius,
                                  float* d_fixed, float* d_moving, double* values, void* stream) {
  using namespace fw::hip;
  try {
    if (!grid) fw::invalid("cc noalloc", "null grid");
    if (!fixed || !moving || !d_fixed || !d_moving || !values) fw::invalid("cc noalloc", "null buffer");
    if (radius == 0 || radius > 8) fw::invalid("cc noalloc", "cc radius " + std::to_string(radius) + " outside 1..=8");
    for (int a = 0; a < 3; ++a)
      if (grid->size[a] <= 0) fw::invalid("grid", "zero size");
    fw::Mat3 dir;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) dir[r][c] = grid->direction[3 * r + c];
    const fw::Grid g({size_t(grid->size[0]), size_t(grid->size[1]), size_t(grid->size[2])},
                     {grid->spacing[0], grid->spacing[1], grid->spacing[2]}, {grid->origin[0], grid->origin[1], grid->origin[2]},
                     dir);
    CcArgs a{};
    a.fixed = fixed;
    a.moving = moving;
    a.valid = nullptr;
    a.d = dims_of(g);
    a.radius = int(radius);
    // the same z runs as fw_cc_pair on a context with the default table, so the same bits
    const uint64_t nv = uint64_t(a.d.n());
    a.tile_z = cc_tile_for(a.d, int(default_tuning().get("cc.tile_z", nv, kCcTileZ)),
                           int(default_tuning().get("cc.min_blocks", nv, 512)));
    const dim3 gr = cc_grid(a.d, a.tile_z);
    const long long blocks = (long long)gr.x * gr.y * gr.z;
    if (blocks > kNoAllocMaxBlocks)
      fw::invalid("cc noalloc", "the grid needs " + std::to_string(blocks) + " CC blocks, more than the " +
                                    std::to_string(kNoAllocMaxBlocks) + " the static scratch holds");
    uint32_t* ptr = scratch_of_current_device();
    a.probes = probes_of(g);
    a.d_fixed = d_fixed;
    a.d_moving = d_moving;
    a.partial_cc = reinterpret_cast<float*>(ptr + kOffCc);
    a.partial_count = reinterpret_cast<int*>(ptr + kOffCount);
    a.bounds = reinterpret_cast<const float*>(ptr + kOffBounds);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = a.d.n();
    float* bound_part = reinterpret_cast<float*>(ptr + kOffBoundPart);
    float* bounds = reinterpret_cast<float*>(ptr + kOffBounds);
    const int bb = int(std::min<long long>(kBoundBlocks, std::max<long long>(1, (n + 255) / 256)));
    launch(k_bound_partials, dim3(bb), dim3(256), 0, s, fixed, moving, n, bound_part);
    launch(k_bound_final, dim3(1), dim3(256), 0, s, (const float*)bound_part, bb, bounds);
    launch_cc_pair(a, s);
    launch(k_cc_value_final, dim3(1), dim3(64), 0, s, (const float*)a.partial_cc, (const int*)a.partial_count, int(blocks),
           values);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("fw_cc_pair_noalloc: kernel launch failed");
    return FW_OK;
  } catch (const fw::Error& e) {
    fw::set_call_error(e.what());
    return FW_ERR_INVALID;
  } catch (const std::exception& e) {
    fw::set_call_error(e.what());
    return FW_ERR_DEVICE;
  }
}
