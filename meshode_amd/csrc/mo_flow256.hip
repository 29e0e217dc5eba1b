// One RK4 step of the config-3 velocity field on FP32 MFMA: the MLP
// 3 -> 256 -> 256 -> 256 -> 3 with softplus, v(y) = MLP(y - 0.5)
// (meshode_amd/neuralode.py's VelocityMLP; the reference's ODEFunc layout,
// neuralode_fast.py:7-33, widened as the config-3 workload specifies), and
// the 3/8-rule step that torchdiffeq's rk4 takes (neuralode_fast.py:55-60):
//   k1 = v(y), k2 = v(y + h k1 / 3), k3 = v(y + h (k2 - k1 / 3)),
//   k4 = v(y + h (k1 - k2 + k3)), y1 = y + h (k1 + 3 (k2 + k3) + k4) / 8.
//
// A workgroup is 8 wavefronts over 64 points; all four stages run in it.
// The 256-wide activations of the 64 points stay in LDS as [unit pair]
// [point][2] (64 KiB per layer, two buffers): the A operand of
// v_mfma_f32_32x32x2f32 (lane l reads point l % 32 of a 32-point block,
// unit 2 kk + l / 32) is then one conflict-free 4-byte read per lane.  Wave
// w computes output units [32 w, 32 w + 32) of a 256 x 256 layer for all
// 64 points: two 32 x 32 accumulator blocks, 128 steps of k = 2; two waves
// share a SIMD, so one's softplus epilogue overlaps the other's MFMAs.
// Its B operand (the weights) comes from global memory (L2-resident,
// 256 KiB per layer) in a packed fragment order in which each lane reads
// one float4 per 4 steps.  The epilogue adds the bias, applies softplus
// (x > 20 ? x : log1p(exp(x)), torch's threshold) and writes the next
// layer's tile.  The 3 -> 256 and 256 -> 3 layers (K = 3, N = 3) run on
// the VALU.
#include <algorithm>

#include "mo_common.h"

#define MO_EXPORT extern "C" __attribute__((visibility("default")))

namespace mo {
namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kH = 256;
constexpr int kPts = 64;
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
// packed parameter layout (floats)
constexpr int kPW0 = 0;                 // W0 [256][4] (k = 3 padded with 0)
constexpr int kPB0 = kPW0 + kH * 4;     // b0 [256]
constexpr int kPW1 = kPB0 + kH;         // W1 fragments [8 col blocks][32 groups][64 lanes][4]
constexpr int kPB1 = kPW1 + kH * kH;
constexpr int kPW2 = kPB1 + kH;
constexpr int kPB2 = kPW2 + kH * kH;
constexpr int kPW3 = kPB2 + kH;         // W3 [3][256]
constexpr int kPB3 = kPW3 + 3 * kH;     // b3 [3] (+1 pad)
constexpr int kPW1T = kPB3 + 4;         // W1^T fragments (backward: dH1 = dA2 W1)
constexpr int kPW2T = kPW1T + kH * kH;  // W2^T fragments (backward: dH2 = dA3 W2)
constexpr int kPacked = kPW2T + kH * kH;

// Activation tile: [128 unit pairs][64 points][2], the point index XOR-ed
// with the unit pair (mod 32) so that the epilogue's writes (32 units, 2
// point rows per wave) spread over the banks; a 32-point block stays a
// permutation of itself, so the MFMA operand reads stay conflict-free.
constexpr int kTile = (kH / 2) * kPts * 2;
__device__ __forceinline__ int tidx(int u, int p) {
  const int kk = u >> 1;
  return kk * (kPts * 2) + 2 * (p ^ (kk & 31)) + (u & 1);
}

// softplus = max(x, 0) + log1p(exp(-|x|)), from the hardware's exp2 and
// log2: the rounding of 1 + exp(-|x|) leaves an absolute error below 1e-7
// (torch computes x > 20 ? x : log1p(exp(x)); beyond 20 both give x to
// within exp(-20)).
__device__ __forceinline__ float softplus(float x) {
  const float e = __builtin_amdgcn_exp2f(-fabsf(x) * 1.4426950408889634f);
  return fmaxf(x, 0.f) + __builtin_amdgcn_logf(1.f + e) * 0.6931471805599453f;
}

// Packs the module's tensors: fragment (cb, g, l, i) of a 256 x 256 layer =
// W[32 cb + l % 32][8 g + 2 i + l / 32] (the B operand of out = in W^T), and
// of its transpose W[8 g + 2 i + l / 32][32 cb + l % 32] (of d_in = d_out W).
__global__ void k_flow256_pack(const float* __restrict__ W0, const float* __restrict__ b0,
                               const float* __restrict__ W1, const float* __restrict__ b1,
                               const float* __restrict__ W2, const float* __restrict__ b2,
                               const float* __restrict__ W3, const float* __restrict__ b3,
                               float* __restrict__ P) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kH * 4) {
    const int u = e >> 2, c = e & 3;
    P[kPW0 + e] = c < 3 ? W0[u * 3 + c] : 0.f;
  }
  if (e < kH) {
    P[kPB0 + e] = b0[e];
    P[kPB1 + e] = b1[e];
    P[kPB2 + e] = b2[e];
  }
  if (e < kH * kH) {
    const int i = e & 3, l = (e >> 2) & 63, g = (e >> 8) & 31, cb = e >> 13;
    const int row = 32 * cb + (l & 31), col = 8 * g + 2 * i + (l >> 5);
    P[kPW1 + e] = W1[row * kH + col];
    P[kPW2 + e] = W2[row * kH + col];
    P[kPW1T + e] = W1[col * kH + row];
    P[kPW2T + e] = W2[col * kH + row];
  }
  if (e < 3 * kH) P[kPW3 + e] = W3[e];
  if (e < 4) P[kPB3 + e] = e < 3 ? b3[e] : 0.f;
}

// acc = Tin W^T for this wave's 32 units [32 w, 32 w + 32) and the 64 (or,
// without kTwo, the first 32) points: 32 x 32 MFMA blocks, 128 steps of
// k = 2.  Result layout: unit 32 w + l % 32, point 32 pb + 8 (r / 4) +
// 4 (l / 32) + r % 4.
// With kStore, the loop also copies Tin to dst (rows [0, nvalid) of 256
// floats, the layout of store_tile), one float4 per thread every fourth
// k-group, so that the copy runs in the shadow of the MFMAs.
// With kStage, it also loads rows [0, nvalid) of src (256 floats each) into
// the tile Tst (one float4 per thread every fourth k-group; Tst must be idle
// during the loop, and a barrier must separate the loop from its readers).
template <bool kTwo, bool kStore = false, bool kStage = false>
__device__ __forceinline__ void mfma_layer(const float* __restrict__ Wf, const float* __restrict__ Tin, int w, int l,
                                           v16f& acc0, v16f& acc1, float* __restrict__ dst = nullptr,
                                           int nvalid = 0, const float* __restrict__ src = nullptr,
                                           float* __restrict__ Tst = nullptr) {
  acc0 = v16f{0};
  acc1 = v16f{0};
  const float4* B = reinterpret_cast<const float4*>(Wf) + (size_t)w * 32 * 64 + l;
  const int hi = l >> 5, pl = l & 31;
  float4 nb[2] = {B[0], B[64]};
#pragma unroll 2
  for (int g = 0; g < 32; ++g) {
    const float4 b = nb[g & 1];
    if (g + 2 < 32) nb[g & 1] = B[(g + 2) * 64];
    float a0[4], a1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int kk = 4 * g + i;
      const int row = kk * (kPts * 2) + hi, sw = kk & 31;
      a0[i] = Tin[row + 2 * (pl ^ sw)];
      a1[i] = kTwo ? Tin[row + 2 * ((32 + pl) ^ sw)] : 0.f;
    }
    const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[i], bb[i], acc0, 0, 0, 0);
      if (kTwo) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[i], bb[i], acc1, 0, 0, 0);
    }
    if (kStore && (g & 3) == 3) {
      const int p = threadIdx.x >> 3, u = 32 * (threadIdx.x & 7) + 4 * (g >> 2);
      if (p < nvalid)
        *reinterpret_cast<float4*>(dst + (size_t)p * kH + u) =
            make_float4(Tin[tidx(u, p)], Tin[tidx(u + 1, p)], Tin[tidx(u + 2, p)], Tin[tidx(u + 3, p)]);
    }
    if (kStage && (g & 3) == 1) {
      const int p = threadIdx.x >> 3, u = 32 * (threadIdx.x & 7) + 4 * (g >> 2);
      if (p < nvalid) {
        const float4 v = *reinterpret_cast<const float4*>(src + (size_t)p * kH + u);
        Tst[tidx(u, p)] = v.x;
        Tst[tidx(u + 1, p)] = v.y;
        Tst[tidx(u + 2, p)] = v.z;
        Tst[tidx(u + 3, p)] = v.w;
      }
    }
  }
}

// Tout = softplus(Tin W^T + b) (forward hidden layer); with kStore Tin is
// also copied to dst (see mfma_layer)
template <bool kTwo, bool kStore = false>
__device__ __forceinline__ void hidden_layer(const float* __restrict__ Wf, const float* __restrict__ bias,
                                             const float* __restrict__ Tin, float* __restrict__ Tout, int w,
                                             int l, float* __restrict__ dst = nullptr, int nvalid = 0) {
  v16f acc0, acc1;
  mfma_layer<kTwo, kStore>(Wf, Tin, w, l, acc0, acc1, dst, nvalid);
  const int u = 32 * w + (l & 31), hi = l >> 5;
  const float bu = bias[u];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int p = 8 * (r >> 2) + 4 * hi + (r & 3);
    Tout[tidx(u, p)] = softplus(acc0[r] + bu);
    if (kTwo) Tout[tidx(u, 32 + p)] = softplus(acc1[r] + bu);
  }
}

// softplus'(a) from h = softplus(a): sigmoid(a) = 1 - exp(-h)
__device__ __forceinline__ float dsoftplus_h(float h) { return 1.f - __builtin_amdgcn_exp2f(-h * 1.4426950408889634f); }

// Tout = (Tin W) * softplus'(A) (backward hidden layer; Wf = W^T fragments):
// H = the layer's saved activations, rows of 256 (the chunk's first row)
template <bool kTwo>
__device__ __forceinline__ void back_layer(const float* __restrict__ WTf, const float* __restrict__ H, int nvalid,
                                           const float* __restrict__ Tin, float* __restrict__ Tout, int w, int l,
                                           float* __restrict__ dst) {
  // the loop also copies Tin to dst and stages the chunk's rows of H into
  // Tout (idle until this epilogue, which reads each value at the position
  // it then overwrites)
  v16f acc0, acc1;
  mfma_layer<kTwo, true, true>(WTf, Tin, w, l, acc0, acc1, dst, nvalid, H, Tout);
  __syncthreads();
  const int j = 32 * w + (l & 31), hi = l >> 5;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int p = 8 * (r >> 2) + 4 * hi + (r & 3);
    Tout[tidx(j, p)] = p < nvalid ? acc0[r] * dsoftplus_h(Tout[tidx(j, p)]) : 0.f;
    if (kTwo) Tout[tidx(j, 32 + p)] = 32 + p < nvalid ? acc1[r] * dsoftplus_h(Tout[tidx(j, 32 + p)]) : 0.f;
  }
}

// dst rows [0, nvalid) of 256 floats <- the tile (coalesced float4 stores)
__device__ __forceinline__ void store_tile(const float* __restrict__ T, float* __restrict__ dst, int nvalid) {
  const int p = threadIdx.x >> 3, part = threadIdx.x & 7;
  if (p >= nvalid) return;
  float4* d = reinterpret_cast<float4*>(dst + (size_t)p * kH + 32 * part);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int u = 32 * part + 4 * q;
    d[q] = make_float4(T[tidx(u, p)], T[tidx(u + 1, p)], T[tidx(u + 2, p)], T[tidx(u + 3, p)]);
  }
}

// Workgroup b of G takes points [n b / G, n (b + 1) / G) in chunks of 64
// (G = one per CU: the shares differ by at most one point, and a chunk of
// at most 32 points skips the second 32-point MFMA block).  With kSave the
// stage inputs x = z - 0.5 go to Xs [4][n][3] and the activations H1, H2, H3
// to Hs [3][4][n][256] (layer, stage, point) for the backward.
template <bool kSave>
__global__ __launch_bounds__(kThreads) void k_flow256_step(const float* __restrict__ y0, int n,
                                                           const float* __restrict__ P, float h,
                                                           float* __restrict__ y1, float* __restrict__ Hs,
                                                           float* __restrict__ Xs) {
  __shared__ __align__(16) float TA[kTile];
  __shared__ __align__(16) float TB[kTile];
  __shared__ float Y[kPts][3], Z[kPts][3], K[4][kPts][3];
  __shared__ float part[kWaves][kPts][3];
  const int tid = threadIdx.x, l = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p0 = (int)((long long)n * blockIdx.x / gridDim.x);
  const int p1 = (int)((long long)n * (blockIdx.x + 1) / gridDim.x);
  const size_t plane = (size_t)n * kH;    // one (layer, stage) plane of Hs
#pragma unroll 1
  for (int base = p0; base < p1; base += kPts) {
    const int nvalid = min(kPts, p1 - base);
    if (tid < kPts) {
      const int i = base + tid;
      for (int c = 0; c < 3; ++c) Y[tid][c] = tid < nvalid ? y0[3 * (size_t)i + c] : 0.5f;
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
      if (tid < kPts) {
        for (int c = 0; c < 3; ++c) {
          const float y = Y[tid][c];
          Z[tid][c] = s == 0 ? y
                    : s == 1 ? y + h * K[0][tid][c] / 3.0f
                    : s == 2 ? y + h * (K[1][tid][c] - K[0][tid][c] / 3.0f)
                             : y + h * (K[0][tid][c] - K[1][tid][c] + K[2][tid][c]);
          if (kSave && tid < nvalid) Xs[((size_t)s * n + base + tid) * 3 + c] = Z[tid][c] - 0.5f;
        }
      }
      __syncthreads();
      // layer 0 (VALU): lane l = point, wave w computes units w + 8 m
      {
        const float x0 = Z[l][0] - 0.5f, x1 = Z[l][1] - 0.5f, x2 = Z[l][2] - 0.5f;
#pragma unroll 4
        for (int u = w; u < kH; u += kWaves) {
          const float4 wr = *reinterpret_cast<const float4*>(P + kPW0 + 4 * u);
          TA[tidx(u, l)] = softplus(P[kPB0 + u] + wr.x * x0 + wr.y * x1 + wr.z * x2);
        }
      }
      __syncthreads();
      {
        float* d1 = kSave ? Hs + (0 * 4 + s) * plane + (size_t)base * kH : nullptr;
        if (nvalid > 32) hidden_layer<true, kSave>(P + kPW1, P + kPB1, TA, TB, w, l, d1, nvalid);
        else hidden_layer<false, kSave>(P + kPW1, P + kPB1, TA, TB, w, l, d1, nvalid);
      }
      __syncthreads();
      {
        float* d2 = kSave ? Hs + (1 * 4 + s) * plane + (size_t)base * kH : nullptr;
        if (nvalid > 32) hidden_layer<true, kSave>(P + kPW2, P + kPB2, TB, TA, w, l, d2, nvalid);
        else hidden_layer<false, kSave>(P + kPW2, P + kPB2, TB, TA, w, l, d2, nvalid);
      }
      __syncthreads();
      if (kSave) store_tile(TA, Hs + (2 * 4 + s) * plane + (size_t)base * kH, nvalid);
      // layer 3 (VALU): wave w sums units [32 w, 32 w + 32) for point l
      {
        float o0 = 0.f, o1 = 0.f, o2 = 0.f;
#pragma unroll 4
        for (int u = 32 * w; u < 32 * w + 32; ++u) {
          const float hv = TA[tidx(u, l)];
          o0 += P[kPW3 + u] * hv;
          o1 += P[kPW3 + kH + u] * hv;
          o2 += P[kPW3 + 2 * kH + u] * hv;
        }
        part[w][l][0] = o0;
        part[w][l][1] = o1;
        part[w][l][2] = o2;
      }
      __syncthreads();
      if (tid < kPts)
        for (int c = 0; c < 3; ++c) {
          float o = 0.f;
          for (int q = 0; q < kWaves; ++q) o += part[q][tid][c];
          K[s][tid][c] = P[kPB3 + c] + o;
        }
      __syncthreads();
    }
    if (tid < nvalid) {
      const int i = base + tid;
      for (int c = 0; c < 3; ++c)
        y1[3 * (size_t)i + c] =
            Y[tid][c] + h * (K[0][tid][c] + 3.0f * (K[1][tid][c] + K[2][tid][c]) + K[3][tid][c]) * 0.125f;
    }
    __syncthreads();
  }
}

// Backward of one step from the saved activations (Hs [3][4][n][256]):
// gy1 = dL/dy1 -> gy0 = dL/dy0; the pre-activation gradients dA1..dA3 go
// to dAs [3][4][n][256] (layer, stage, point) and the output gradients
// dL/dk_s to Gs [4][n][3], from which the caller forms the parameter
// gradients (sums over stages and points).  Per stage, last first:
// dA3 = (W3^T g) * softplus'(A3) (VALU), dA2 = (dA3 W2) * softplus'(A2),
// dA1 = (dA2 W1) * softplus'(A1) (MFMA), dz = W0^T dA1 (VALU), and the
// RK4 chain: y1 = y + h (k1 + 3 k2 + 3 k3 + k4) / 8, z2 = y + h k1 / 3,
// z3 = y + h (k2 - k1 / 3), z4 = y + h (k1 - k2 + k3).
__global__ __launch_bounds__(kThreads) void k_flow256_step_bwd(int n, const float* __restrict__ P, float h,
                                                               const float* __restrict__ Hs,
                                                               const float* __restrict__ gy1,
                                                               float* __restrict__ gy0, float* __restrict__ dAs,
                                                               float* __restrict__ Gs) {
  __shared__ __align__(16) float TA[kTile];
  __shared__ __align__(16) float TB[kTile];
  __shared__ float DK[4][kPts][3], DY[kPts][3];
  __shared__ float part[kWaves][kPts][3];
  const int tid = threadIdx.x, l = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p0 = (int)((long long)n * blockIdx.x / gridDim.x);
  const int p1 = (int)((long long)n * (blockIdx.x + 1) / gridDim.x);
  const size_t plane = (size_t)n * kH;
  const float third = h / 3.0f;
#pragma unroll 1
  for (int base = p0; base < p1; base += kPts) {
    const int nvalid = min(kPts, p1 - base);
    if (tid < kPts)
      for (int c = 0; c < 3; ++c) {
        const float g = tid < nvalid ? gy1[3 * (size_t)(base + tid) + c] : 0.f;
        DY[tid][c] = g;
        const float a = h * g * 0.125f;
        DK[0][tid][c] = a;
        DK[1][tid][c] = 3.0f * a;
        DK[2][tid][c] = 3.0f * a;
        DK[3][tid][c] = a;
      }
    __syncthreads();
#pragma unroll 1
    for (int s = 3; s >= 0; --s) {
      const float* H1 = Hs + (0 * 4 + s) * plane + (size_t)base * kH;
      const float* H2 = Hs + (1 * 4 + s) * plane + (size_t)base * kH;
      const float* H3 = Hs + (2 * 4 + s) * plane + (size_t)base * kH;
      if (tid < nvalid)
        for (int c = 0; c < 3; ++c) Gs[((size_t)s * n + base + tid) * 3 + c] = DK[s][tid][c];
      // dA3 (VALU): lane l = point, wave w = units [32 w, 32 w + 32)
      {
        const float g0 = DK[s][l][0], g1 = DK[s][l][1], g2 = DK[s][l][2];
        const bool ok = l < nvalid;
#pragma unroll 4
        for (int q = 0; q < 8; ++q) {
          const int u = 32 * w + 4 * q;
          const float4 hv = ok ? *reinterpret_cast<const float4*>(H3 + (size_t)l * kH + u) : make_float4(0, 0, 0, 0);
          const float hs[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float dh = P[kPW3 + u + e] * g0 + P[kPW3 + kH + u + e] * g1 + P[kPW3 + 2 * kH + u + e] * g2;
            TA[tidx(u + e, l)] = ok ? dh * dsoftplus_h(hs[e]) : 0.f;
          }
        }
      }
      __syncthreads();
      {
        float* d3 = dAs + (2 * 4 + s) * plane + (size_t)base * kH;
        if (nvalid > 32) back_layer<true>(P + kPW2T, H2, nvalid, TA, TB, w, l, d3);
        else back_layer<false>(P + kPW2T, H2, nvalid, TA, TB, w, l, d3);
      }
      __syncthreads();
      {
        float* d2 = dAs + (1 * 4 + s) * plane + (size_t)base * kH;
        if (nvalid > 32) back_layer<true>(P + kPW1T, H1, nvalid, TB, TA, w, l, d2);
        else back_layer<false>(P + kPW1T, H1, nvalid, TB, TA, w, l, d2);
      }
      __syncthreads();
      store_tile(TA, dAs + (0 * 4 + s) * plane + (size_t)base * kH, nvalid);
      // dz = W0^T dA1 (VALU): wave w sums units [32 w, 32 w + 32) for point l
      {
        float o0 = 0.f, o1 = 0.f, o2 = 0.f;
#pragma unroll 4
        for (int u = 32 * w; u < 32 * w + 32; ++u) {
          const float d = TA[tidx(u, l)];
          const float4 wr = *reinterpret_cast<const float4*>(P + kPW0 + 4 * u);
          o0 += wr.x * d;
          o1 += wr.y * d;
          o2 += wr.z * d;
        }
        part[w][l][0] = o0;
        part[w][l][1] = o1;
        part[w][l][2] = o2;
      }
      __syncthreads();
      if (tid < kPts)
        for (int c = 0; c < 3; ++c) {
          float dz = 0.f;
          for (int q = 0; q < kWaves; ++q) dz += part[q][tid][c];
          DY[tid][c] += dz;
          if (s == 3) {          // z4 = y + h (k1 - k2 + k3)
            DK[0][tid][c] += h * dz;
            DK[1][tid][c] -= h * dz;
            DK[2][tid][c] += h * dz;
          } else if (s == 2) {   // z3 = y + h (k2 - k1 / 3)
            DK[1][tid][c] += h * dz;
            DK[0][tid][c] -= third * dz;
          } else if (s == 1) {   // z2 = y + h k1 / 3
            DK[0][tid][c] += third * dz;
          }
        }
      __syncthreads();
    }
    if (tid < nvalid)
      for (int c = 0; c < 3; ++c) gy0[3 * (size_t)(base + tid) + c] = DY[tid][c];
    __syncthreads();
  }
}

// The small parameter gradients of a step from the saved tensors (rows =
// the 4n stage points): dW0 = dA1^T X [256][3], db0, db1, db2 (column sums of
// dA1, dA2, dA3), dW3 = G^T H3 [3][256], db3 (column sums of G).  Thread u of
// workgroup b sums rows [R b / G, R (b + 1) / G) for unit u (coalesced rows
// of dA1, dA2, dA3, H3) and writes the workgroup's partial row;
// k_flow256_rows_reduce adds the rows in a fixed order (FP64).
constexpr int kSmallGrad = 3 * kH + 3 * kH + 3 * kH + 3;   // 2307
__global__ __launch_bounds__(kH) void k_flow256_grad_small(int rows, const float* __restrict__ dA1,
                                                           const float* __restrict__ dA2,
                                                           const float* __restrict__ dA3,
                                                           const float* __restrict__ H3, const float* __restrict__ X,
                                                           const float* __restrict__ G, float* __restrict__ partial) {
  const int u = threadIdx.x;
  const int r0 = (int)((long long)rows * blockIdx.x / gridDim.x);
  const int r1 = (int)((long long)rows * (blockIdx.x + 1) / gridDim.x);
  float w00 = 0.f, w01 = 0.f, w02 = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f, w30 = 0.f, w31 = 0.f, w32 = 0.f;
  float g0s = 0.f, g1s = 0.f, g2s = 0.f;
#pragma unroll 4
  for (int r = r0; r < r1; ++r) {
    const size_t o = (size_t)r * kH + u;
    const float a1 = dA1[o], a2 = dA2[o], a3 = dA3[o], h3 = H3[o];
    const float x0 = X[3 * (size_t)r], x1 = X[3 * (size_t)r + 1], x2 = X[3 * (size_t)r + 2];
    const float g0 = G[3 * (size_t)r], g1 = G[3 * (size_t)r + 1], g2 = G[3 * (size_t)r + 2];
    w00 += a1 * x0;
    w01 += a1 * x1;
    w02 += a1 * x2;
    b0 += a1;
    b1 += a2;
    b2 += a3;
    w30 += g0 * h3;
    w31 += g1 * h3;
    w32 += g2 * h3;
    g0s += g0;
    g1s += g1;
    g2s += g2;
  }
  float* row = partial + (size_t)blockIdx.x * kSmallGrad;
  row[3 * u] = w00;
  row[3 * u + 1] = w01;
  row[3 * u + 2] = w02;
  row[3 * kH + u] = b0;
  row[4 * kH + u] = b1;
  row[5 * kH + u] = b2;
  row[6 * kH + u] = w30;
  row[7 * kH + u] = w31;
  row[8 * kH + u] = w32;
  if (u == 0) {
    row[9 * kH] = g0s;
    row[9 * kH + 1] = g1s;
    row[9 * kH + 2] = g2s;
  }
}

// out[e] = sum over the nblocks rows (stride `len`) of partial[b][e], in a
// fixed order (FP64): a block covers 64 elements, wave w adds rows w, w + 8,
// ..., and the eight wave sums are added in wave order
__global__ __launch_bounds__(512) void k_flow256_rows_reduce(const float* __restrict__ partial, int nblocks, int len,
                                                             float* __restrict__ out) {
  __shared__ double S[8][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  const float* p = partial + (size_t)blockIdx.y * nblocks * len;
  double s = 0.0;
  if (e < len) {
#pragma unroll 4
    for (int b = w; b < nblocks; b += 8) s += p[(size_t)b * len + e];
  }
  S[w][lane] = s;
  __syncthreads();
  if (w == 0 && e < len) {
    double t = S[0][lane];
    for (int k = 1; k < 8; ++k) t += S[k][lane];
    out[(size_t)blockIdx.y * len + e] = (float)t;
  }
}

// dW1 = dA2^T H1 and dW2 = dA3^T H2 over the 4n stage points of a step
// (blockIdx.y picks the pair): split over the rows, workgroup b of G takes
// rows [R b / G, R (b + 1) / G) in chunks of 32, staged through LDS (rows
// padded to 288 floats: the two half-waves' reads fall in disjoint banks)
// with the next chunk prefetched into registers.  Wave w accumulates
// dW[32 w + i][j] for all 256 j: eight 32 x 32 MFMA blocks (A = dA rows
// transposed, B = H rows, k = 2 rows per step).  Partials per workgroup;
// k_flow256_rows_reduce adds them in a fixed order (FP64).
constexpr int kDwRows = 32;
constexpr int kDwLd = 288;
__global__ __launch_bounds__(kThreads) void k_flow256_dw(int rows, const float* __restrict__ dA2,
                                                         const float* __restrict__ H1, const float* __restrict__ dA3,
                                                         const float* __restrict__ H2, float* __restrict__ partial) {
  __shared__ __align__(16) float SD[kDwRows * kDwLd];
  __shared__ __align__(16) float SH[kDwRows * kDwLd];
  const float* D = blockIdx.y ? dA3 : dA2;
  const float* Hm = blockIdx.y ? H2 : H1;
  const int tid = threadIdx.x, l = tid & 63, hi = l >> 5, pl = l & 31;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r0 = (int)((long long)rows * blockIdx.x / gridDim.x);
  const int r1 = (int)((long long)rows * (blockIdx.x + 1) / gridDim.x);
  v16f acc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) acc[c] = v16f{0};
  float4 pd[4], ph[4];
  auto fetch = [&](int base) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kThreads * q, row = idx >> 6, c4 = idx & 63;
      const bool ok = base + row < r1;
      pd[q] = ok ? *reinterpret_cast<const float4*>(D + (size_t)(base + row) * kH + 4 * c4) : make_float4(0, 0, 0, 0);
      ph[q] = ok ? *reinterpret_cast<const float4*>(Hm + (size_t)(base + row) * kH + 4 * c4) : make_float4(0, 0, 0, 0);
    }
  };
  if (r0 < r1) fetch(r0);
#pragma unroll 1
  for (int base = r0; base < r1; base += kDwRows) {
    __syncthreads();                       // the previous chunk's reads are done
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kThreads * q, row = idx >> 6, c4 = idx & 63;
      *reinterpret_cast<float4*>(SD + row * kDwLd + 4 * c4) = pd[q];
      *reinterpret_cast<float4*>(SH + row * kDwLd + 4 * c4) = ph[q];
    }
    __syncthreads();
    if (base + kDwRows < r1) fetch(base + kDwRows);
#pragma unroll 4
    for (int kk = 0; kk < kDwRows / 2; ++kk) {
      const int row = (2 * kk + hi) * kDwLd;
      const float a = SD[row + 32 * w + pl];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, SH[row + 32 * c + pl], acc[c], 0, 0, 0);
    }
  }
  float* out = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (kH * kH);
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int u = 32 * w + 8 * (r >> 2) + 4 * hi + (r & 3);
      out[u * kH + 32 * c + pl] = acc[c][r];
    }
}

int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

}  // namespace
}  // namespace mo

using namespace mo;

MO_EXPORT int64_t mo_flow256_packed_floats(void) { return kPacked; }

MO_EXPORT int mo_flow256_pack(const float* W0, const float* b0, const float* W1, const float* b1, const float* W2,
                              const float* b2, const float* W3, const float* b3, float* packed, void* stream) {
  MO_CHECK(W0 && b0 && W1 && b1 && W2 && b2 && W3 && b3 && packed, MO_ERR_ARG, "mo_flow256_pack: null buffer");
  hipLaunchKernelGGL(k_flow256_pack, dim3(kH * kH / 256), dim3(256), 0, (hipStream_t)stream, W0, b0, W1, b1, W2, b2,
                     W3, b3, packed);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int mo_flow256_step(const float* y0, int64_t n, const float* packed, float h, float* y1, void* stream) {
  MO_CHECK(packed && n >= 0 && n <= (1LL << 30) && (n == 0 || (y0 && y1)), MO_ERR_ARG,
           "mo_flow256_step: bad arguments");
  if (n == 0) return MO_OK;
  const int grid = (int)std::min<int64_t>((n + kPts - 1) / kPts, cu_count());
  hipLaunchKernelGGL(k_flow256_step<false>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, y0, (int)n, packed, h,
                     y1, nullptr, nullptr);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int mo_flow256_step_save(const float* y0, int64_t n, const float* packed, float h, float* y1, float* Hs,
                                   float* Xs, void* stream) {
  MO_CHECK(packed && n >= 0 && n <= (1LL << 30) / 12 && (n == 0 || (y0 && y1 && Hs && Xs)), MO_ERR_ARG,
           "mo_flow256_step_save: bad arguments");
  MO_CHECK(((uintptr_t)Hs & 15) == 0, MO_ERR_ARG, "mo_flow256_step_save: Hs not 16-byte aligned");
  if (n == 0) return MO_OK;
  const int grid = (int)std::min<int64_t>((n + kPts - 1) / kPts, cu_count());
  hipLaunchKernelGGL(k_flow256_step<true>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, y0, (int)n, packed, h,
                     y1, Hs, Xs);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int mo_flow256_step_backward(int64_t n, const float* packed, float h, const float* Hs, const float* gy1,
                                       float* gy0, float* dAs, float* Gs, void* stream) {
  MO_CHECK(packed && n >= 0 && n <= (1LL << 30) / 12 && (n == 0 || (Hs && gy1 && gy0 && dAs && Gs)), MO_ERR_ARG,
           "mo_flow256_step_backward: bad arguments");
  MO_CHECK(((uintptr_t)Hs & 15) == 0 && ((uintptr_t)dAs & 15) == 0, MO_ERR_ARG,
           "mo_flow256_step_backward: Hs / dAs not 16-byte aligned");
  if (n == 0) return MO_OK;
  const int grid = (int)std::min<int64_t>((n + kPts - 1) / kPts, cu_count());
  hipLaunchKernelGGL(k_flow256_step_bwd, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, (int)n, packed, h, Hs,
                     gy1, gy0, dAs, Gs);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int64_t mo_flow256_small_grad_floats(void) { return kSmallGrad; }

// out [2307]: dW0 [256][3], db0 [256], db1 [256], db2 [256], dW3 [3][256], db3 [3]
MO_EXPORT int mo_flow256_small_grads(int64_t n, const float* dAs, const float* Hs, const float* Xs, const float* Gs,
                                     float* out, float* work, void* stream) {
  MO_CHECK(n > 0 && n <= (1LL << 30) / 12 && dAs && Hs && Xs && Gs && out && work, MO_ERR_ARG,
           "mo_flow256_small_grads: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const size_t plane4 = (size_t)4 * n * kH;          // one layer, all four stages
  const int rows = (int)(4 * n);
  const int nb = 256;
  hipLaunchKernelGGL(k_flow256_grad_small, dim3(nb), dim3(kH), 0, s, rows, dAs, dAs + plane4, dAs + 2 * plane4,
                     Hs + 2 * plane4, Xs, Gs, work);
  hipLaunchKernelGGL(k_flow256_rows_reduce, dim3((kSmallGrad + 63) / 64, 1), dim3(512), 0, s, work, nb, kSmallGrad,
                     out);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int64_t mo_flow256_dw_work_floats(void) { return (int64_t)2 * cu_count() * kH * kH; }

// out [2][256][256] = dW1, dW2 of a step (rows = the 4n stage points of
// mo_flow256_step_backward's dAs and mo_flow256_step_save's Hs); work:
// mo_flow256_dw_work_floats() floats (device)
MO_EXPORT int mo_flow256_dw_grads(int64_t n, const float* dAs, const float* Hs, float* out, float* work,
                                  void* stream) {
  MO_CHECK(n > 0 && n <= (1LL << 30) / 12 && dAs && Hs && out && work, MO_ERR_ARG, "mo_flow256_dw_grads: bad arguments");
  MO_CHECK(((uintptr_t)dAs & 15) == 0 && ((uintptr_t)Hs & 15) == 0, MO_ERR_ARG,
           "mo_flow256_dw_grads: dAs / Hs not 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const size_t plane4 = (size_t)4 * n * kH;
  const int G = cu_count();
  hipLaunchKernelGGL(k_flow256_dw, dim3(G, 2), dim3(kThreads), 0, s, (int)(4 * n), dAs + plane4, Hs,
                     dAs + 2 * plane4, Hs + plane4, work);
  hipLaunchKernelGGL(k_flow256_rows_reduce, dim3(kH * kH / 64, 2), dim3(512), 0, s, work, G, kH * kH, out);
  MO_HIP(hipGetLastError());
  return MO_OK;
}
