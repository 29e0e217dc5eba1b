// Device-side source topology for the LM solver: per-vertex adjacency lists
// with edge multiplicities, built from the residual blocks of
// Deformer::Deform (one EdgeLoss per half-edge (F[i][j], F[i][(j+1)%3]),
// deformer.cc:40-53) or Deformer::DeformGraph (one per graph edge,
// deformer.cc:394-403), for a whole batch of pairs at once.
//
//   1. k_topo_count   : every edge (a != b) counts +1 at a and at b.
//   2. scan           : exclusive prefix sum of the counts -> list offsets.
//   3. k_topo_fill    : every edge appends the other endpoint to both lists.
//   4. k_topo_finish  : each vertex sorts its (short) list, merges equal
//                       neighbours into (neighbour, multiplicity) entries,
//                       and records (offset, unique count) and its degree.
// The result does not depend on the order of the atomic appends.
#include "mo_common.h"
#include "mo_topo.h"

namespace mo {

namespace {

__device__ inline bool edge_of(const TopoPairDesc& P, long long h, int* a, int* b) {
  if (P.kind == 0) {
    long long f = h / 3;
    int j = (int)(h - f * 3);
    *a = P.idx[f * 3 + j];
    *b = P.idx[f * 3 + (j == 2 ? 0 : j + 1)];
  } else {
    *a = P.idx[h * 2];
    *b = P.idx[h * 2 + 1];
  }
  return *a != *b;
}

__global__ void k_zero_u32(unsigned* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_topo_count(const TopoPairDesc* __restrict__ pairs, unsigned* __restrict__ cnt,
                             int* __restrict__ err) {
  const TopoPairDesc P = pairs[blockIdx.y];
  for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < P.nedges;
       h += (long long)gridDim.x * blockDim.x) {
    int a, b;
    bool ok = edge_of(P, h, &a, &b);
    if (a < 0 || b < 0 || a >= P.nv || b >= P.nv) { atomicOr(err, 1); continue; }
    if (!ok) continue;
    atomicAdd(&cnt[P.vbase + a], 1u);
    atomicAdd(&cnt[P.vbase + b], 1u);
  }
}

__global__ void k_topo_fill(const TopoPairDesc* __restrict__ pairs, const long long* __restrict__ off,
                            unsigned* __restrict__ cursor, unsigned* __restrict__ raw) {
  const TopoPairDesc P = pairs[blockIdx.y];
  for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < P.nedges;
       h += (long long)gridDim.x * blockDim.x) {
    int a, b;
    bool ok = edge_of(P, h, &a, &b);
    if (!ok || a < 0 || b < 0 || a >= P.nv || b >= P.nv) continue;
    unsigned pa = atomicAdd(&cursor[P.vbase + a], 1u);
    unsigned pb = atomicAdd(&cursor[P.vbase + b], 1u);
    raw[off[P.vbase + a] + pa] = (unsigned)b;
    raw[off[P.vbase + b] + pb] = (unsigned)a;
  }
}

__global__ void k_topo_finish(const TopoPairDesc* __restrict__ pairs, const long long* __restrict__ off,
                              const unsigned* __restrict__ cnt, unsigned* __restrict__ raw, int2* __restrict__ adj,
                              float* __restrict__ deg, int* __restrict__ pstat, int* __restrict__ err) {
  const int p = blockIdx.y;
  const TopoPairDesc P = pairs[p];
  int uniq = 0, maxm = 0;
  for (long long lv = (long long)blockIdx.x * blockDim.x + threadIdx.x; lv < P.nv;
       lv += (long long)gridDim.x * blockDim.x) {
    const long long v = P.vbase + lv;
    const long long o = off[v];
    const int n = (int)cnt[v];
    unsigned* L = raw + o;
    // insertion sort (lists are short: the vertex degree)
    for (int i = 1; i < n; ++i) {
      unsigned x = L[i];
      int j = i - 1;
      while (j >= 0 && L[j] > x) { L[j + 1] = L[j]; --j; }
      L[j + 1] = x;
    }
    int u = 0;
    for (int i = 0; i < n;) {
      unsigned x = L[i];
      int m = 0;
      while (i < n && L[i] == x) { ++m; ++i; }
      if (m > 255 || x > 0xFFFFFFu) { atomicOr(err, 2); m = m > 255 ? 255 : m; }
      maxm = m > maxm ? m : maxm;
      L[u++] = (x & 0xFFFFFFu) | ((unsigned)m << 24);
    }
    adj[v] = make_int2((int)o, u);
    deg[v] = (float)n;
    uniq += u;
  }
  if (uniq) atomicAdd(&pstat[2 * p], uniq);
  if (maxm) atomicMax(&pstat[2 * p + 1], maxm);
}

// Exclusive scan of unsigned counts into 64-bit offsets: per-block scan,
// block totals scanned by one block, then added back.  Workgroups of 256
// threads (one wavefront per SIMD) fit beside a running grid build.
constexpr int kScanT = 256;

__global__ void __launch_bounds__(kScanT) k_scan_blocks(const unsigned* __restrict__ in, long long n,
                                                        long long* __restrict__ out, long long* __restrict__ bsum) {
  __shared__ long long s[kScanT];
  long long i = (long long)blockIdx.x * kScanT + threadIdx.x;
  long long x = i < n ? (long long)in[i] : 0;
  s[threadIdx.x] = x;
  __syncthreads();
  for (int d = 1; d < kScanT; d <<= 1) {
    long long t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - x;
  if (threadIdx.x == kScanT - 1) bsum[blockIdx.x] = s[kScanT - 1];
}

__global__ void __launch_bounds__(kScanT) k_scan_sums(long long* __restrict__ bsum, long long nb,
                                                      long long* __restrict__ total) {
  __shared__ long long s[kScanT];
  long long carry = 0;
  for (long long base = 0; base < nb; base += kScanT) {
    long long i = base + threadIdx.x;
    long long x = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int d = 1; d < kScanT; d <<= 1) {
      long long t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - x;
    long long tot = s[kScanT - 1];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_add(long long* __restrict__ out, long long n, const long long* __restrict__ bsum) {
  long long i = (long long)blockIdx.x * kScanT + threadIdx.x;
  if (i < n) out[i] += bsum[blockIdx.x];
}

}  // namespace

size_t scan_workspace_bytes(long long n) { return (size_t)(((n + kScanT - 1) / kScanT) + 1) * sizeof(long long); }

int exclusive_scan_launch(const unsigned* in, long long n, long long* out, long long* total_dev, void* ws,
                          hipStream_t s) {
  if (n <= 0) return MO_OK;
  long long nb = (n + kScanT - 1) / kScanT;
  long long* bsum = (long long*)ws;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanT), 0, s, in, n, out, bsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanT), 0, s, bsum, nb, total_dev);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanT), 0, s, out, n, bsum);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

size_t topo_workspace_bytes(long long nvtot) {
  return (size_t)nvtot * (2 * sizeof(unsigned) + sizeof(long long)) + scan_workspace_bytes(nvtot) + 64;
}

int topo_build_launch(const TopoPairDesc* pairs_dev, int npairs, long long max_edges, long long max_nv,
                      long long nvtot, unsigned* raw, int2* adj, float* deg, int* pstat, void* ws, int* err_dev,
                      hipStream_t s) {
  if (npairs <= 0 || nvtot <= 0) return MO_OK;
  char* w = (char*)ws;
  unsigned* cnt = (unsigned*)w; w += (size_t)nvtot * sizeof(unsigned);
  unsigned* cursor = (unsigned*)w; w += (size_t)nvtot * sizeof(unsigned);
  long long* off = (long long*)w; w += (size_t)nvtot * sizeof(long long);
  void* scan_ws = (void*)w;
  {
    long long nz = nvtot * 2;   // cnt and cursor (adjacent)
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, cnt, nz);
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((2 * npairs + 255) / 256)), dim3(256), 0, s, (unsigned*)pstat,
                       (long long)(2 * npairs));
  }
  const int T = 256;
  long long bx = (max_edges + T - 1) / T;
  if (bx > 4096) bx = 4096;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL(k_topo_count, dim3((unsigned)bx, (unsigned)npairs), dim3(T), 0, s, pairs_dev, cnt, err_dev);
  int rc = exclusive_scan_launch(cnt, nvtot, off, nullptr, scan_ws, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_topo_fill, dim3((unsigned)bx, (unsigned)npairs), dim3(T), 0, s, pairs_dev, off, cursor, raw);
  long long bv = (max_nv + T - 1) / T;
  if (bv > 4096) bv = 4096;
  if (bv < 1) bv = 1;
  hipLaunchKernelGGL(k_topo_finish, dim3((unsigned)bv, (unsigned)npairs), dim3(T), 0, s, pairs_dev, off, cnt, raw,
                     adj, deg, pstat, err_dev);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

}  // namespace mo
