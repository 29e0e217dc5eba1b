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
constexpr int kPacked = kPB3 + 4;

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
// W[32 cb + l % 32][8 g + 2 i + l / 32].
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
  }
  if (e < 3 * kH) P[kPW3 + e] = W3[e];
  if (e < 4) P[kPB3 + e] = e < 3 ? b3[e] : 0.f;
}

// Tout = softplus(Tin W^T + b) for this wave's 32 units [32 w, 32 w + 32)
// and all 64 points (two 32 x 32 MFMA blocks, 128 steps of k = 2)
template <bool kTwo>
__device__ __forceinline__ void hidden_layer(const float* __restrict__ Wf, const float* __restrict__ bias,
                                             const float* __restrict__ Tin, float* __restrict__ Tout, int w,
                                             int l) {
  v16f acc0 = v16f{0}, acc1 = v16f{0};
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
  }
  // D layout: unit 32 w + l % 32, point 32 pb + 8 (r / 4) + 4 (l / 32) + r % 4
  const int u = 32 * w + pl;
  const float bu = bias[u];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int p = 8 * (r >> 2) + 4 * hi + (r & 3);
    Tout[tidx(u, p)] = softplus(acc0[r] + bu);
    if (kTwo) Tout[tidx(u, 32 + p)] = softplus(acc1[r] + bu);
  }
}

// Workgroup b of G takes points [n b / G, n (b + 1) / G) in chunks of 64
// (G = one per CU: the shares differ by at most one point, and a chunk of
// at most 32 points skips the second 32-point MFMA block).
__global__ __launch_bounds__(kThreads) void k_flow256_step(const float* __restrict__ y0, int n,
                                                           const float* __restrict__ P, float h,
                                                           float* __restrict__ y1) {
  __shared__ __align__(16) float TA[kTile];
  __shared__ __align__(16) float TB[kTile];
  __shared__ float Y[kPts][3], Z[kPts][3], K[4][kPts][3];
  __shared__ float part[kWaves][kPts][3];
  const int tid = threadIdx.x, l = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p0 = (int)((long long)n * blockIdx.x / gridDim.x);
  const int p1 = (int)((long long)n * (blockIdx.x + 1) / gridDim.x);
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
      if (nvalid > 32) hidden_layer<true>(P + kPW1, P + kPB1, TA, TB, w, l);
      else hidden_layer<false>(P + kPW1, P + kPB1, TA, TB, w, l);
      __syncthreads();
      if (nvalid > 32) hidden_layer<true>(P + kPW2, P + kPB2, TB, TA, w, l);
      else hidden_layer<false>(P + kPW2, P + kPB2, TB, TA, w, l);
      __syncthreads();
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
  hipLaunchKernelGGL(k_flow256_step, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, y0, (int)n, packed, h, y1);
  MO_HIP(hipGetLastError());
  return MO_OK;
}
