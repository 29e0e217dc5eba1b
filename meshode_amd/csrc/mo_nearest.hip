// Nearest source point of every target point: the k = 1 query of
// ReverseLossLayer (reverse_loss_layer.py:15-26, scipy cKDTree(src).query(tar)).
//
// Brute force on the device.  Squared distances are evaluated in FP64 from
// the FP32 coordinates as (dx*dx + dy*dy) + dz*dz without contraction, so a
// CPU evaluation in the same order gives the same bits; among equal distances
// the lowest source index wins (cKDTree leaves ties unspecified).
//
// Layout: sources are split into `chunks` contiguous ranges, one per
// blockIdx.y, so that a few thousand targets still spread over the whole
// chip; each block stages its range through LDS in tiles of kTile points
// (FP64, 24 KiB) and every thread keeps the best (d2, index) of one target.
// A second kernel reduces the per-chunk winners in chunk order.
#include <algorithm>

#include "mo_common.h"

#define MO_EXPORT extern "C" __attribute__((visibility("default")))

namespace mo {
namespace {

constexpr int kThreads = 256;
constexpr int kTile = 1024;

__global__ __launch_bounds__(kThreads) void k_nearest_chunk(const float* __restrict__ src, int ns,
                                                           const float* __restrict__ tar, int nt, int chunk,
                                                           double* __restrict__ best_d, int* __restrict__ best_i) {
#pragma clang fp contract(off)
  __shared__ double sx[kTile], sy[kTile], sz[kTile];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  const int lo = blockIdx.y * chunk;
  const int hi = min(ns, lo + chunk);
  double px = 0, py = 0, pz = 0;
  if (t < nt) {
    px = (double)tar[3 * (size_t)t];
    py = (double)tar[3 * (size_t)t + 1];
    pz = (double)tar[3 * (size_t)t + 2];
  }
  double bd = __builtin_inf();
  int bi = -1;
  for (int base = lo; base < hi; base += kTile) {
    const int m = min(kTile, hi - base);
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += kThreads) {
      const size_t g = 3 * (size_t)(base + k);
      sx[k] = (double)src[g];
      sy[k] = (double)src[g + 1];
      sz[k] = (double)src[g + 2];
    }
    __syncthreads();
    for (int k = 0; k < m; ++k) {
      const double dx = sx[k] - px, dy = sy[k] - py, dz = sz[k] - pz;
      const double d = (dx * dx + dy * dy) + dz * dz;
      if (d < bd) {
        bd = d;
        bi = base + k;
      }
    }
  }
  if (t < nt) {
    best_d[(size_t)blockIdx.y * nt + t] = bd;
    best_i[(size_t)blockIdx.y * nt + t] = bi;
  }
}

__global__ void k_nearest_reduce(int nt, int chunks, const double* __restrict__ best_d, const int* __restrict__ best_i,
                                 int* __restrict__ index, double* __restrict__ dist2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double bd = __builtin_inf();
  int bi = -1;
  for (int c = 0; c < chunks; ++c) {
    const double d = best_d[(size_t)c * nt + t];
    const int i = best_i[(size_t)c * nt + t];
    if (i >= 0 && (bi < 0 || d < bd)) {
      bd = d;
      bi = i;
    }
  }
  index[t] = bi;
  if (dist2) dist2[t] = bd;
}

}  // namespace

int nearest_f32(const float* src, int64_t ns, const float* tar, int64_t nt, int32_t* index, double* dist2,
                hipStream_t s) {
  MO_CHECK(src && ns > 0 && nt >= 0 && ns < (1LL << 31) && nt < (1LL << 31) && (nt == 0 || (tar && index)),
           MO_ERR_ARG, "nearest: bad arguments (ns %lld, nt %lld)", (long long)ns, (long long)nt);
  if (nt == 0) return MO_OK;
  const int bx = (int)((nt + kThreads - 1) / kThreads);
  // enough blocks for the 256 CUs, chunks of at least one tile
  int chunks = std::max(1, std::min((int)((ns + kTile - 1) / kTile), (2048 + bx - 1) / bx));
  chunks = std::min(chunks, 65535);
  const int chunk = (int)((ns + chunks - 1) / chunks);
  chunks = (int)((ns + chunk - 1) / chunk);
  void* ws = nullptr;
  const size_t nd = (size_t)chunks * (size_t)nt;
  MO_HIP(hipMallocAsync(&ws, nd * 12 + 256, s));
  double* bd = (double*)ws;
  int* bi = (int*)((char*)ws + ((nd * 8 + 255) & ~size_t(255)));
  hipLaunchKernelGGL(k_nearest_chunk, dim3(bx, chunks), dim3(kThreads), 0, s, src, (int)ns, tar, (int)nt, chunk, bd,
                     bi);
  hipLaunchKernelGGL(k_nearest_reduce, dim3(bx), dim3(kThreads), 0, s, (int)nt, chunks, bd, bi, index, dist2);
  const hipError_t e = hipGetLastError();
  (void)hipFreeAsync(ws, s);
  MO_CHECK(e == hipSuccess, MO_ERR_HIP, "nearest: launch failed: %s", hipGetErrorString(e));
  return MO_OK;
}

}  // namespace mo

MO_EXPORT int mo_nearest_f32(const float* src, int64_t ns, const float* tar, int64_t nt, int32_t* index,
                             double* dist2, void* stream) {
  return mo::nearest_f32(src, ns, tar, nt, index, dist2, (hipStream_t)stream);
}
