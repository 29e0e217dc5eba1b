// pyDeform layer kernels: the per-vertex distance loss and the
// edge-rigidity losses with their hand-written backward passes.
//
//   NormalizeByTemplate / DenormalizeByTemplate   normalize.cc:5-45
//   DistanceFieldLoss_forward / _backward          distance_layer.cc:8-80
//   StoreRigidityInformation / RigidEdgeLoss_*     rigid_layer.cc:11-133
//   StoreGraphInformation / GraphEdgeLoss_*        graph_layer.cc:11-135
//
// Compiled with -ffp-contract=off: results are bit-identical to the
// reference's FP32 loops.  The backward scatter of the reference
// (sequential `g[a] -= r; g[b] += r` in half-edge order) is re-expressed as
// a per-vertex gather over an incidence list sorted by (half-edge, role),
// which performs the same additions in the same order for every vertex —
// bit-exact and free of atomics.
#include "mo_common.h"
#include "mo_layers.h"
#include "mo_lookup.h"

#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>

namespace mo {

namespace {

__global__ void k_normalize_by_template(float* __restrict__ V, long long n3, double scale, double t0, double t1,
                                        double t2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  int a = (int)(i % 3);
  double t = a == 0 ? t0 : (a == 1 ? t1 : t2);
  V[i] = (float)(((double)V[i] - t) / scale);
}

__global__ void k_denormalize_by_template(float* __restrict__ V, long long n3, double scale, double t0, double t1,
                                          double t2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  int a = (int)(i % 3);
  double t = a == 0 ? t0 : (a == 1 ? t1 : t2);
  V[i] = (float)((double)V[i] * scale + t);
}

__global__ void k_distance_fwd(const float* __restrict__ grid, int N, const float* __restrict__ V, long long n,
                               float* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float d = lookup_f32(grid, N, V[i * 3], V[i * 3 + 1], V[i * 3 + 2]);
  out[i] = d * d;
}

__global__ void k_distance_bwd(const float* __restrict__ grid, int N, const float* __restrict__ V, long long n,
                               float* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  JetF d = lookup_jet(grid, N, V[i * 3], V[i * 3 + 1], V[i * 3 + 2]);
  // vd = d * d (Jet product): vd.v = d.a*d.v + d.v*d.a; out = vd.v * 0.5
  float x0 = d.a * d.v0, y0 = d.v0 * d.a;
  float x1 = d.a * d.v1, y1 = d.v1 * d.a;
  float x2 = d.a * d.v2, y2 = d.v2 * d.a;
  out[i * 3 + 0] = (float)((double)(x0 + y0) * 0.5);
  out[i * 3 + 1] = (float)((double)(x1 + y1) * 0.5);
  out[i * 3 + 2] = (float)((double)(x2 + y2) * 0.5);
}

// Directed edge h of the two layouts: faces (h = 3f + j, a = F[f][j],
// b = F[f][(j+1)%3]) or an explicit [E,2] list (a = E[h][0], b = E[h][1]).
__device__ inline void edge_ends(const int* __restrict__ idx, int kind, long long h, int* a, int* b) {
  if (kind == 0) {
    long long f = h / 3;
    int j = (int)(h - f * 3);
    *a = idx[f * 3 + j];
    *b = idx[f * 3 + (j == 2 ? 0 : j + 1)];
  } else {
    *a = idx[h * 2];
    *b = idx[h * 2 + 1];
  }
}

__global__ void k_edge_store(const float* __restrict__ V, long long nv, const int* __restrict__ idx, int kind,
                             long long nh, float* __restrict__ off, int* __restrict__ err) {
  long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  int a, b;
  edge_ends(idx, kind, h, &a, &b);
  if (a < 0 || b < 0 || a >= nv || b >= nv) {
    atomicOr(err, 1);
    off[h * 3] = off[h * 3 + 1] = off[h * 3 + 2] = 0.0f;
    return;
  }
  for (int k = 0; k < 3; ++k) off[h * 3 + k] = V[(long long)b * 3 + k] - V[(long long)a * 3 + k];
}

__global__ void k_edge_fwd(const float* __restrict__ V, long long nv, const int* __restrict__ idx, int kind,
                           long long nh, const float* __restrict__ off, float* __restrict__ out,
                           int* __restrict__ err) {
  long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  int a, b;
  edge_ends(idx, kind, h, &a, &b);
  if (a < 0 || b < 0 || a >= nv || b >= nv) {   // err may be null: the indices were checked on the host
    if (err) atomicOr(err, 1);
    out[h * 3] = out[h * 3 + 1] = out[h * 3 + 2] = 0.0f;
    return;
  }
  for (int k = 0; k < 3; ++k) {
    float l = V[(long long)b * 3 + k] - V[(long long)a * 3 + k] - off[h * 3 + k];
    out[h * 3 + k] = l * l;
  }
}

// Per-vertex gather over the sorted incidence list: entry = (h << 1) | role,
// role 0 = tail a (subtract), role 1 = head b (add).
__global__ void k_edge_bwd(const float* __restrict__ V, long long nv, const int* __restrict__ idx, int kind,
                           const float* __restrict__ off, const long long* __restrict__ rp,
                           const long long* __restrict__ inc, float* __restrict__ out) {
  long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nv) return;
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
  for (long long e = rp[u]; e < rp[u + 1]; ++e) {
    long long code = inc[e];
    long long h = code >> 1;
    int a, b;
    edge_ends(idx, kind, h, &a, &b);
    if (a < 0 || b < 0 || a >= nv || b >= nv) continue;   // index buffer changed since the lists were built
    float r0 = V[(long long)b * 3 + 0] - V[(long long)a * 3 + 0] - off[h * 3 + 0];
    float r1 = V[(long long)b * 3 + 1] - V[(long long)a * 3 + 1] - off[h * 3 + 1];
    float r2 = V[(long long)b * 3 + 2] - V[(long long)a * 3 + 2] - off[h * 3 + 2];
    if (code & 1) { g0 = g0 + r0; g1 = g1 + r1; g2 = g2 + r2; }
    else { g0 = g0 - r0; g1 = g1 - r1; g2 = g2 - r2; }
  }
  out[u * 3 + 0] = g0;
  out[u * 3 + 1] = g1;
  out[u * 3 + 2] = g2;
}

// Incidence items of the 2 nh directed-edge ends: item q = 2 h + role is
// keyed by its vertex (tail for role 0, head for role 1); an end outside
// [0, nv) gets the key nv, which sorts after every vertex and is left out
// of the lists (and flagged).
__global__ void k_inc_items(const int* __restrict__ idx, int kind, long long nh, long long nv,
                            unsigned* __restrict__ key, long long* __restrict__ val, int* __restrict__ err) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * nh) return;
  int a, b;
  edge_ends(idx, kind, q >> 1, &a, &b);
  const int v = (q & 1) ? b : a;
  const bool ok = a >= 0 && b >= 0 && a < nv && b < nv;
  if (!ok) atomicOr(err, 1);
  key[q] = ok ? (unsigned)v : (unsigned)nv;
  val[q] = q;
}

// rp[v] = first position of key v in the sorted keys (v = 0..nv).
__global__ void k_inc_rows(const unsigned* __restrict__ key, long long n, long long nv, long long* __restrict__ rp) {
  long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nv) return;
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((long long)key[mid] < v) lo = mid + 1; else hi = mid;
  }
  rp[v] = lo;
}

inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

inline int key_bits(long long nv) {
  int b = 1;
  while (b < 32 && (1LL << b) <= nv) ++b;
  return b;
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int normalize_by_template_launch(float* V, long long nv, double scale, const double* trans, bool inverse,
                                 hipStream_t s) {
  if (nv <= 0) return MO_OK;
  long long n3 = nv * 3;
  if (!inverse)
    hipLaunchKernelGGL(k_normalize_by_template, dim3(blocks_for(n3, 256)), dim3(256), 0, s, V, n3, scale, trans[0],
                       trans[1], trans[2]);
  else
    hipLaunchKernelGGL(k_denormalize_by_template, dim3(blocks_for(n3, 256)), dim3(256), 0, s, V, n3, scale,
                       trans[0], trans[1], trans[2]);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

int distance_loss_launch(const float* grid, int N, const float* V, long long nv, float* out, bool backward,
                         hipStream_t s) {
  if (nv <= 0) return MO_OK;
  if (!backward)
    hipLaunchKernelGGL(k_distance_fwd, dim3(blocks_for(nv, 256)), dim3(256), 0, s, grid, N, V, nv, out);
  else
    hipLaunchKernelGGL(k_distance_bwd, dim3(blocks_for(nv, 256)), dim3(256), 0, s, grid, N, V, nv, out);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

int edge_store_launch(const float* V, long long nv, const int* idx, int kind, long long nh, float* off, int* err,
                      hipStream_t s) {
  if (nh <= 0) return MO_OK;
  hipLaunchKernelGGL(k_edge_store, dim3(blocks_for(nh, 256)), dim3(256), 0, s, V, nv, idx, kind, nh, off, err);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

int edge_fwd_launch(const float* V, long long nv, const int* idx, int kind, long long nh, const float* off,
                    float* out, int* err, hipStream_t s) {
  if (nh <= 0) return MO_OK;
  hipLaunchKernelGGL(k_edge_fwd, dim3(blocks_for(nh, 256)), dim3(256), 0, s, V, nv, idx, kind, nh, off, out, err);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

size_t edge_incidence_workspace_bytes(long long nh, long long nv) {
  const long long n = std::max<long long>(2 * nh, 1);
  size_t temp = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                           (const long long*)nullptr, (long long*)nullptr, n, 0, key_bits(nv));
  return 2 * al256((size_t)n * sizeof(unsigned)) + 2 * al256((size_t)n * sizeof(long long)) +
         al256((size_t)(nv + 1) * sizeof(long long)) + al256(temp) + 256;
}

int edge_incidence_build(const int* idx, int kind, long long nh, long long nv, void* ws, long long** rp,
                         long long** inc, int* err, hipStream_t s) {
  const long long n = std::max<long long>(2 * nh, 1);
  char* w = (char*)ws;
  unsigned* k_in = (unsigned*)w; w += al256((size_t)n * sizeof(unsigned));
  unsigned* k_out = (unsigned*)w; w += al256((size_t)n * sizeof(unsigned));
  long long* v_in = (long long*)w; w += al256((size_t)n * sizeof(long long));
  long long* v_out = (long long*)w; w += al256((size_t)n * sizeof(long long));
  long long* r = (long long*)w; w += al256((size_t)(nv + 1) * sizeof(long long));
  size_t temp = 0;
  MO_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, k_in, k_out, v_in, v_out, n, 0, key_bits(nv), s));
  if (nh > 0) {
    hipLaunchKernelGGL(k_inc_items, dim3(blocks_for(2 * nh, 256)), dim3(256), 0, s, idx, kind, nh, nv, k_in, v_in,
                       err);
    MO_HIP(hipGetLastError());
    // LSD radix sort is stable: each vertex's items stay in increasing q =
    // 2 h + role, the order of the reference's sequential loop
    MO_HIP(hipcub::DeviceRadixSort::SortPairs((void*)w, temp, k_in, k_out, v_in, v_out, 2 * nh, 0, key_bits(nv), s));
  }
  hipLaunchKernelGGL(k_inc_rows, dim3(blocks_for(nv + 1, 256)), dim3(256), 0, s, k_out, nh > 0 ? 2 * nh : 0, nv, r);
  MO_HIP(hipGetLastError());
  *rp = r;
  *inc = v_out;
  return MO_OK;
}

int edge_bwd_launch(const float* V, long long nv, const int* idx, int kind, const float* off, const long long* rp,
                    const long long* inc, float* out, hipStream_t s) {
  if (nv <= 0) return MO_OK;
  hipLaunchKernelGGL(k_edge_bwd, dim3(blocks_for(nv, 256)), dim3(256), 0, s, V, nv, idx, kind, off, rp, inc, out);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

}  // namespace mo
