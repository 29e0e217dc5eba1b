// Nearest source point of every target point: the k = 1 query of
// ReverseLossLayer (reverse_loss_layer.py:15-26, scipy cKDTree(src).query(tar)).
//
// Squared distances are evaluated in FP64 from the FP32 coordinates as
// (dx*dx + dy*dy) + dz*dz without contraction, so a CPU evaluation in the
// same order gives the same bits; among equal distances the lowest source
// index wins (cKDTree leaves ties unspecified).
//
// Two exact searches with the same result:
//
// * brute force (fewer than kGridMin sources): sources are split into
//   `chunks` contiguous ranges, one per blockIdx.y, so that a few thousand
//   targets still spread over the whole chip; each block stages its range
//   through LDS in tiles of kTile points (FP64, 24 KiB) and every thread
//   keeps the best (d2, index) of one target.  A second kernel reduces the
//   per-chunk winners in chunk order.
// * a uniform grid (kGridMin sources or more): the sources' bounding box is
//   cut into cubic cells (at most kGridMax per axis, about 2 sources per cell
//   of the box), the sources are counting-sorted by cell (x fastest), and
//   each target (eight lanes) visits the cells in shells of growing
//   Chebyshev radius around its own cell, one contiguous range of sorted
//   sources per row of cells.  It stops once every unvisited cell lies farther than the best
//   distance found (an FP64 bound with a margin of 1e-12 of the box size, so
//   points exactly as far are still visited).  The winner is the least
//   (d2, index), which does not depend on the order of visits: the same
//   index and distance bits as the brute force.
#include <algorithm>
#include <cmath>

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

// ---- uniform grid ----------------------------------------------------------
constexpr int kGridMin = 4096;   // sources from which the grid search is used
constexpr int kGridMax = 64;     // cells per axis at most

struct NNGrid {
  double lo[3];
  double h, inv_h, margin;
  int dim[3];
  int ncells;
};

// (d, i) replaces (bd, bi) when less; nothing replaces a start of
// (inf, -1) except a smaller distance, as in the brute force
__device__ inline void better(double d, int i, double& bd, int& bi) {
  if (d < bd || (d == bd && i < bi)) {
    bd = d;
    bi = i;
  }
}

// bounding box of the sources and the grid over it (one block)
__global__ __launch_bounds__(1024) void k_nn_box(const float* __restrict__ src, int ns, int gmax,
                                                 NNGrid* __restrict__ grid) {
  __shared__ float red[6][1024];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < ns; i += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      const float v = src[3 * (size_t)i + a];
      mn[a] = fminf(mn[a], v);
      mx[a] = fmaxf(mx[a], v);
    }
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = mn[a];
    red[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    NNGrid g;
    double ext = 0.0, mag = 0.0;
    for (int a = 0; a < 3; ++a) {
      g.lo[a] = (double)red[a][0];
      ext = fmax(ext, (double)red[3 + a][0] - (double)red[a][0]);
      mag = fmax(mag, fmax(fabs((double)red[a][0]), fabs((double)red[3 + a][0])));
    }
    g.h = ext > 0.0 ? ext / gmax : 1.0;
    g.inv_h = 1.0 / g.h;
    g.margin = 1e-12 * (ext + mag);
    g.ncells = 1;
    for (int a = 0; a < 3; ++a) {
      const double n = ceil(((double)red[3 + a][0] - (double)red[a][0]) * g.inv_h);
      g.dim[a] = (int)fmin(fmax(n, 1.0), (double)gmax);
      g.ncells *= g.dim[a];
    }
    *grid = g;
  }
}

__device__ inline int cell_coord(double v, double lo, double inv_h, int dim) {
  const double c = floor((v - lo) * inv_h);
  return (int)fmin(fmax(c, 0.0), (double)(dim - 1));
}

__global__ void k_nn_zero(int* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

__global__ void k_nn_count(const float* __restrict__ src, int ns, const NNGrid* __restrict__ grid,
                           int* __restrict__ cell_of, int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const NNGrid g = *grid;
  int c[3];
  for (int a = 0; a < 3; ++a) c[a] = cell_coord((double)src[3 * (size_t)i + a], g.lo[a], g.inv_h, g.dim[a]);
  const int cell = (c[2] * g.dim[1] + c[1]) * g.dim[0] + c[0];
  cell_of[i] = cell;
  atomicAdd(&count[cell], 1);
}

// Exclusive scan of the cell counts in two passes: k_nn_scan_local scans
// each run of 1024 cells (one block) and records the run's total;
// k_nn_scan_add adds the totals of the preceding runs, copies the starts to
// the scatter cursors and closes the array with start[ncells] = ns.
constexpr int kScan = 1024;
__global__ __launch_bounds__(kScan) void k_nn_scan_local(const NNGrid* __restrict__ grid,
                                                         const int* __restrict__ count, int* __restrict__ start,
                                                         int* __restrict__ run_total) {
  __shared__ int part[kScan];
  const int n = grid->ncells;
  const int c = blockIdx.x * kScan + threadIdx.x;
  if ((int)(blockIdx.x * kScan) >= n) return;   // whole block: no barrier is skipped by part of it
  const int v = c < n ? count[c] : 0;
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = 1; w < kScan; w <<= 1) {   // inclusive scan
    const int u = (int)threadIdx.x >= w ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += u;
    __syncthreads();
  }
  if (c < n) start[c] = part[threadIdx.x] - v;
  if (threadIdx.x == kScan - 1) run_total[blockIdx.x] = part[kScan - 1];
}

__global__ __launch_bounds__(256) void k_nn_scan_add(const NNGrid* __restrict__ grid, const int* __restrict__ run_total,
                                                     int ns, int* __restrict__ start, int* __restrict__ cursor) {
  __shared__ int red[256];
  const int n = grid->ncells;
  const int run = blockIdx.x / (kScan / 256);
  if (run * kScan >= n) return;
  int s = 0;
  for (int r = threadIdx.x; r < run; r += 256) s += run_total[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < n) {
    const int v = start[c] + red[0];
    start[c] = v;
    cursor[c] = v;
  }
  if (c == 0) start[n] = ns;
}

__global__ void k_nn_scatter(const float* __restrict__ src, int ns, const int* __restrict__ cell_of,
                             int* __restrict__ cursor, float4* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int pos = atomicAdd(&cursor[cell_of[i]], 1);
  sorted[pos] = make_float4(src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2], __int_as_float(i));
}

// Eight lanes per target: the group reads the source ranges of up to eight
// rows of cells at once (one row per lane), scans each range with the
// lanes striding over its sources, and reduces its (d2, index) by shuffles
// after every shell.
constexpr int kGroup = 8;

__device__ inline void scan_strided(const float4* __restrict__ sorted, int b, int e, int j, const double p[3],
                                    double& bd, int& bi) {
#pragma clang fp contract(off)
  for (int k = b + j; k < e; k += kGroup) {
    const float4 q = sorted[k];
    const double dx = (double)q.x - p[0], dy = (double)q.y - p[1], dz = (double)q.z - p[2];
    better((dx * dx + dy * dy) + dz * dz, __float_as_int(q.w), bd, bi);
  }
}

__global__ __launch_bounds__(256) void k_nn_query(const float* __restrict__ tar, int nt,
                                                  const NNGrid* __restrict__ grid, const int* __restrict__ start,
                                                  const float4* __restrict__ sorted, int* __restrict__ index,
                                                  double* __restrict__ dist2) {
#pragma clang fp contract(off)
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = gt / kGroup, j = gt % kGroup;
  if (t >= nt) return;                    // all eight lanes of the group together
  const NNGrid g = *grid;
  const double p[3] = {(double)tar[3 * (size_t)t], (double)tar[3 * (size_t)t + 1], (double)tar[3 * (size_t)t + 2]};
  int qc[3];
  for (int a = 0; a < 3; ++a) qc[a] = cell_coord(p[a], g.lo[a], g.inv_h, g.dim[a]);
  const int kmax = max(g.dim[0], max(g.dim[1], g.dim[2]));
  double bd = INFINITY;
  int bi = -1;
  for (int k = 0; k < kmax; ++k) {
    // shell k: the cells of the box qc +- k not in the box qc +- (k - 1),
    // clipped to the grid; rows (z, y) on the box's edge are whole x ranges,
    // the others its two end cells
    const int z0 = max(qc[2] - k, 0), z1 = min(qc[2] + k, g.dim[2] - 1);
    const int y0 = max(qc[1] - k, 0), y1 = min(qc[1] + k, g.dim[1] - 1);
    const int x0 = max(qc[0] - k, 0), x1 = min(qc[0] + k, g.dim[0] - 1);
    const int ny = y1 - y0 + 1, nrows = (z1 - z0 + 1) * ny;
    for (int rb = 0; rb < nrows; rb += kGroup) {
      int bA = 0, eA = 0, bB = 0, eB = 0;
      const int r = rb + j;
      if (r < nrows) {
        const int z = z0 + r / ny, y = y0 + r % ny;
        const int row = (z * g.dim[1] + y) * g.dim[0];
        if (z == qc[2] - k || z == qc[2] + k || y == qc[1] - k || y == qc[1] + k) {
          bA = start[row + x0];
          eA = start[row + x1 + 1];
        } else {
          if (qc[0] - k >= 0) {
            bA = start[row + qc[0] - k];
            eA = start[row + qc[0] - k + 1];
          }
          if (k > 0 && qc[0] + k < g.dim[0]) {
            bB = start[row + qc[0] + k];
            eB = start[row + qc[0] + k + 1];
          }
        }
      }
      const int m1 = min(kGroup, nrows - rb);
      for (int m = 0; m < m1; ++m) {
        scan_strided(sorted, __shfl(bA, m, kGroup), __shfl(eA, m, kGroup), j, p, bd, bi);
        scan_strided(sorted, __shfl(bB, m, kGroup), __shfl(eB, m, kGroup), j, p, bd, bi);
      }
    }
    for (int o = kGroup / 2; o > 0; o >>= 1) {
      const double od = __shfl_xor(bd, o, kGroup);
      const int oi = __shfl_xor(bi, o, kGroup);
      if (oi >= 0) better(od, oi, bd, bi);
    }
    // every unvisited source lies beyond a face of the visited box that is
    // not on the grid's boundary: the nearest such face bounds its distance
    double bound = INFINITY;
    for (int a = 0; a < 3; ++a) {
      if (qc[a] - k > 0) bound = fmin(bound, p[a] - (g.lo[a] + (qc[a] - k) * g.h));
      if (qc[a] + k < g.dim[a] - 1) bound = fmin(bound, (g.lo[a] + (qc[a] + k + 1) * g.h) - p[a]);
    }
    if (bound == INFINITY) break;          // the whole grid has been visited
    const double b = bound - g.margin;
    if (b > 0.0 && bd < b * b) break;
  }
  if (j == 0) {
    index[t] = bi;
    if (dist2) dist2[t] = bd;
  }
}

int nearest_grid(const float* src, int ns, const float* tar, int nt, int32_t* index, double* dist2, hipStream_t s) {
  const int gmax = std::max(1, std::min(kGridMax, (int)std::lround(std::cbrt(2.0 * ns))));
  const size_t cells = (size_t)gmax * gmax * gmax;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const int runs = (int)((cells + kScan - 1) / kScan);
  const size_t bytes = al(sizeof(NNGrid)) + al((cells + 1) * 4) * 3 + al((size_t)runs * 4) + al((size_t)ns * 4) +
                       al((size_t)ns * 16);
  char* ws = nullptr;
  MO_HIP(hipMallocAsync((void**)&ws, bytes, s));
  char* p = ws;
  NNGrid* grid = (NNGrid*)p;
  p += al(sizeof(NNGrid));
  int* count = (int*)p;
  p += al((cells + 1) * 4);
  int* start = (int*)p;
  p += al((cells + 1) * 4);
  int* cursor = (int*)p;
  p += al((cells + 1) * 4);
  int* run_total = (int*)p;
  p += al((size_t)runs * 4);
  int* cell_of = (int*)p;
  p += al((size_t)ns * 4);
  float4* sorted = (float4*)p;
  const int nb = (ns + 255) / 256;
  hipLaunchKernelGGL(k_nn_box, dim3(1), dim3(1024), 0, s, src, ns, gmax, grid);
  hipLaunchKernelGGL(k_nn_zero, dim3((unsigned)std::min<size_t>((cells + 255) / 256, 1024)), dim3(256), 0, s, count,
                     (int)cells);
  hipLaunchKernelGGL(k_nn_count, dim3(nb), dim3(256), 0, s, src, ns, grid, cell_of, count);
  hipLaunchKernelGGL(k_nn_scan_local, dim3(runs), dim3(kScan), 0, s, grid, count, start, run_total);
  hipLaunchKernelGGL(k_nn_scan_add, dim3(runs * (kScan / 256)), dim3(256), 0, s, grid, run_total, ns, start, cursor);
  hipLaunchKernelGGL(k_nn_scatter, dim3(nb), dim3(256), 0, s, src, ns, cell_of, cursor, sorted);
  hipLaunchKernelGGL(k_nn_query, dim3((unsigned)(((size_t)nt * kGroup + 255) / 256)), dim3(256), 0, s, tar, nt, grid,
                     start, sorted, index, dist2);
  const hipError_t e = hipGetLastError();
  (void)hipFreeAsync(ws, s);
  MO_CHECK(e == hipSuccess, MO_ERR_HIP, "nearest: launch failed: %s", hipGetErrorString(e));
  return MO_OK;
}

}  // namespace

int nearest_f32(const float* src, int64_t ns, const float* tar, int64_t nt, int32_t* index, double* dist2,
                hipStream_t s) {
  MO_CHECK(src && ns > 0 && nt >= 0 && ns < (1LL << 31) && nt < (1LL << 31) && (nt == 0 || (tar && index)),
           MO_ERR_ARG, "nearest: bad arguments (ns %lld, nt %lld)", (long long)ns, (long long)nt);
  if (nt == 0) return MO_OK;
  if (ns >= kGridMin) return nearest_grid(src, (int)ns, tar, (int)nt, index, dist2, s);
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
