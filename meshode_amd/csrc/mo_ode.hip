// One fixed-grid RK4 step of the NeuralODE flow (src/python/layers/
// neuralode_fast.py:7-64: ODEFunc 4 -> 50 -> 50 -> 50 -> 3 with LeakyReLU
// 0.01, odeint(method="rk4") on a two-point grid) fused into two kernels:
//
//   forward : y1 = y0 + (k1 + 3 (k2 + k3) + k4) dt / 8 with the 3/8-rule
//             stages k_s = sign * MLP(cat(z_s, tau_s) - 0.5),
//             z1 = y0, z2 = y0 + dt k1 / 3, z3 = y0 + dt (k2 - k1 / 3),
//             z4 = y0 + dt (k1 - k2 + k3);
//   backward: dL/dy1 -> dL/dy0 and dL/dparameters through the four stages.
//
// tau_s (the network's time input), sign and dt come from the host: a
// decreasing grid is integrated as t -> -t with f -> -f(-t, y) (sign -1).
//
// A block is 4 wavefronts over 64 points: lane t of every wave works on
// point t, and wave w computes the units [13 w, 13 w + 13) of each layer.
// The parameters (W0 50x4, b0, W1 50x50, b1, W2 50x50, b2, W3 3x50, b3,
// row-major, the module's parameter order) are copied to LDS with rows
// padded to 16-byte alignment; a weight row is read by all lanes at one
// address (broadcast) against the point's 50 inputs held in registers.  A
// layer's outputs go to an LDS tile [unit][point] from which every wave
// reloads the next layer's inputs.
//
// The backward recomputes each stage's activations from the saved stage
// inputs and walks the layers down, keeping two such tiles: per layer the
// block's weight and bias gradients are sums over its points of products
// of two tile rows (float4 reads along the points, FP32 over the chunk's 64
// points), added in FP64 into the block's own row of partials in global
// memory, and the deltas one layer down use weight copies cut into the
// waves' unit blocks (one delta read per input row).
// Blocks stride over the 64-point chunks (at most kMaxBlocks of them); a
// last kernel adds the blocks' rows in block order in FP64, so results do
// not depend on scheduling and the sums over many points keep FP32
// accuracy.
#include <algorithm>

#include "mo_common.h"

#define MO_EXPORT extern "C" __attribute__((visibility("default")))

namespace mo {
namespace {

constexpr int kH = 50;
// parameter offsets in the module's flat order
constexpr int kOW0 = 0, kOB0 = 200, kOW1 = 250, kOB1 = 2750, kOW2 = 2800, kOB2 = 5300, kOW3 = 5350, kOB3 = 5500;
constexpr int kNP = 5503;
// LDS layout (floats): rows of 50 padded to 52 (16-byte aligned rows)
constexpr int kRS = 52;
constexpr int kLW0 = 0, kLB0 = 200, kLW1 = 252, kLB1 = 2852, kLW2 = 2904, kLB2 = 5504, kLW3 = 5556, kLB3 = 5712;
constexpr int kFwdP = 5716;
// backward only: W1 and W2 with each row cut into the waves' unit blocks,
// block w at 16 w (13 units, 3 zero slots): row r = 64 floats
constexpr int kBS = 64;
constexpr int kLW1B = 5716, kLW2B = kLW1B + kH * kBS;
constexpr int kBwdP = kLW2B + kH * kBS;

constexpr int kPts = 64;           // points per block
constexpr int kWaves = 4;
constexpr int kThreads = kPts * kWaves;
constexpr int kUW = 13;            // units per wave: wave w has [13 w, min(13 w + 13, 50))
constexpr int kLD = 64;            // tile row stride (floats): one row per unit, one column per point
constexpr int kTile = kH * kLD;
constexpr int kMaxBlocks = 512;    // backward grid cap (bounds the partials)

__device__ __forceinline__ float lrelu(float a) { return a > 0.f ? a : a * 0.01f; }
// derivative from the activation: lrelu(a) > 0 exactly when a > 0 (torch's
// LeakyReLU backward takes the slope 0.01 at a = 0)
__device__ __forceinline__ float dlrelu(float h) { return h > 0.f ? 1.f : 0.01f; }

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// The same LDS array behind an offset the compiler cannot see through: the
// parameter reads of one stage are then not hoisted out of the stage loop
// (which would pin the output layer's weights in registers).
__device__ __forceinline__ const float* opaque_base(const float* P) {
  int o = 0;
  asm volatile("" : "+v"(o));
  return P + o;
}

// params (flat) -> LDS layout; the wave-blocked W1, W2 copies too when `blocked`
__device__ __forceinline__ void load_params(const float* __restrict__ params, float* __restrict__ P, bool blocked) {
  for (int e = threadIdx.x; e < kNP; e += blockDim.x) {
    const float v = params[e];
    if (e < kOB0) {
      P[kLW0 + e] = v;
    } else if (e < kOW1) {
      P[kLB0 + e - kOB0] = v;
    } else if (e < kOB1) {
      const int q = e - kOW1, r = q / kH, c = q - kH * r;
      P[kLW1 + kRS * r + c] = v;
    } else if (e < kOW2) {
      P[kLB1 + e - kOB1] = v;
    } else if (e < kOB2) {
      const int q = e - kOW2, r = q / kH, c = q - kH * r;
      P[kLW2 + kRS * r + c] = v;
    } else if (e < kOW3) {
      P[kLB2 + e - kOB2] = v;
    } else if (e < kOB3) {
      const int q = e - kOW3, r = q / kH, c = q - kH * r;
      P[kLW3 + kRS * r + c] = v;
    } else {
      P[kLB3 + e - kOB3] = v;
    }
  }
  if (blocked) {
    for (int e = threadIdx.x; e < kH * kBS; e += blockDim.x) {
      const int r = e / kBS, slot = e % kBS, m = slot & 15, c = kUW * (slot >> 4) + m;
      const bool used = m < kUW && c < kH;
      P[kLW1B + e] = used ? params[kOW1 + kH * r + c] : 0.f;
      P[kLW2B + e] = used ? params[kOW2 + kH * r + c] : 0.f;
    }
  }
}

// w[0:50] . h (w 16-byte aligned)
__device__ __forceinline__ float dot50(const float* __restrict__ w, const float (&h)[kH]) {
  float a = 0.f;
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    const float4 u = *reinterpret_cast<const float4*>(w + 4 * q);
    a += u.x * h[4 * q] + u.y * h[4 * q + 1] + u.z * h[4 * q + 2] + u.w * h[4 * q + 3];
  }
  const float2 u = *reinterpret_cast<const float2*>(w + 48);
  return a + u.x * h[48] + u.y * h[49];
}

// h[j] = T[j][t]
__device__ __forceinline__ void load_col(const float* __restrict__ T, int t, float (&h)[kH]) {
#pragma unroll
  for (int j = 0; j < kH; ++j) h[j] = T[j * kLD + t];
}

__device__ __forceinline__ int units_end(int w) { return min(kUW * w + kUW, kH); }

// this wave's units of layer 0: T[i][t] = lrelu(b0[i] + W0[i] . x)
__device__ __forceinline__ void layer0_rows(const float* __restrict__ P, const float x[4], float* __restrict__ T, int w,
                                            int t) {
#pragma unroll 1
  for (int i = kUW * w; i < units_end(w); ++i) {
    const float4 u = *reinterpret_cast<const float4*>(P + kLW0 + 4 * i);
    T[i * kLD + t] = lrelu(P[kLB0 + i] + u.x * x[0] + u.y * x[1] + u.z * x[2] + u.w * x[3]);
  }
}

// this wave's units of a hidden 50 -> 50 layer
__device__ __forceinline__ void hidden_rows(const float* __restrict__ W, const float* __restrict__ b,
                                            const float (&h)[kH], float* __restrict__ T, int w, int t) {
#pragma unroll 1
  for (int i = kUW * w; i < units_end(w); ++i) T[i * kLD + t] = lrelu(b[i] + dot50(W + kRS * i, h));
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_ode_fwd(
    const float* __restrict__ y0, int n, const float* __restrict__ params, float4 tau4, float sign, float dt,
    float* __restrict__ y1, float* __restrict__ zsave) {
  __shared__ __align__(256) float T1[kTile];
  __shared__ __align__(256) float T2[kTile];
  __shared__ __align__(16) float P[kFwdP];
  load_params(params, P, false);
  const int w = wave_id(), t = threadIdx.x & (kPts - 1);
  const int i = blockIdx.x * kPts + t;
  const bool valid = i < n;
  const float third = 1.0f / 3.0f;
  float y[3] = {0.5f, 0.5f, 0.5f};
  if (valid)
    for (int c = 0; c < 3; ++c) y[c] = y0[3 * i + c];
  const float taus[4] = {tau4.x, tau4.y, tau4.z, tau4.w};
  float k1[3] = {0.f, 0.f, 0.f}, k2[3] = {0.f, 0.f, 0.f}, k3[3] = {0.f, 0.f, 0.f}, k4[3];
  __syncthreads();
#pragma unroll 1
  for (int s = 0; s < 4; ++s) {
    float z[3];
    for (int c = 0; c < 3; ++c)
      z[c] = s == 0 ? y[c]
           : s == 1 ? y[c] + dt * k1[c] * third
           : s == 2 ? y[c] + dt * (k2[c] - k1[c] * third)
                    : y[c] + dt * (k1[c] - k2[c] + k3[c]);
    if (s > 0 && zsave && valid && w == 0)
      for (int c = 0; c < 3; ++c) zsave[9 * i + 3 * (s - 1) + c] = z[c];
    const float x[4] = {z[0] - 0.5f, z[1] - 0.5f, z[2] - 0.5f, taus[s] - 0.5f};
    const float* Ps = opaque_base(P);
    float h[kH];
    layer0_rows(Ps, x, T1, w, t);
    __syncthreads();
    load_col(T1, t, h);
    hidden_rows(Ps + kLW1, Ps + kLB1, h, T2, w, t);
    __syncthreads();
    load_col(T2, t, h);
    hidden_rows(Ps + kLW2, Ps + kLB2, h, T1, w, t);
    __syncthreads();
    load_col(T1, t, h);
    float k[3];
    for (int c = 0; c < 3; ++c) k[c] = sign * (Ps[kLB3 + c] + dot50(Ps + kLW3 + kRS * c, h));
    for (int c = 0; c < 3; ++c) {
      if (s == 0) k1[c] = k[c];
      else if (s == 1) k2[c] = k[c];
      else if (s == 2) k3[c] = k[c];
      else k4[c] = k[c];
    }
    __syncthreads();
  }
  if (valid && w == 0)
    for (int c = 0; c < 3; ++c) y1[3 * i + c] = y[c] + (k1[c] + 3.0f * (k2[c] + k3[c]) + k4[c]) * dt * 0.125f;
}

// Sums over the block's 64 points of tile rows (256-byte aligned rows of
// 64 floats).  Lane l visits the row's 16 float4 chunks in the order
// q ^ (l & 15), so the 16 lanes of a group read 16 different bank quads
// whatever rows they read; the per-lane order is fixed, so results are
// reproducible.
__device__ __forceinline__ float dot_pts(const float* __restrict__ A, const float* __restrict__ B) {
  const int rot = threadIdx.x & 15;
  float s = 0.f;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int c = 4 * (q ^ rot);
    const float4 u = *reinterpret_cast<const float4*>(A + c);
    const float4 v = *reinterpret_cast<const float4*>(B + c);
    s += u.x * v.x + u.y * v.y + u.z * v.z + u.w * v.w;
  }
  return s;
}

__device__ __forceinline__ float sum_pts(const float* __restrict__ A) {
  const int rot = threadIdx.x & 15;
  float s = 0.f;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const float4 u = *reinterpret_cast<const float4*>(A + 4 * (q ^ rot));
    s += (u.x + u.y) + (u.z + u.w);
  }
  return s;
}

// The chunk's gradient of a layer out = W in + b (W: R x C): dW[r][c] =
// sum_p D[r][p] In[c][p], db[r] = sum_p D[r][p], added into the block's row
// of partials (FP64, written on its first pass).  Each element stays with
// one thread, which alone reads and writes it.
__device__ __forceinline__ void accum_layer(const float* __restrict__ D, const float* __restrict__ In, int R, int C,
                                            int w_off, int b_off, double* __restrict__ row, bool first) {
  for (int e = threadIdx.x; e < R * C; e += kThreads) {
    const double v = dot_pts(D + (e / C) * kLD, In + (e % C) * kLD);
    row[w_off + e] = first ? v : row[w_off + e] + v;
  }
  for (int r = threadIdx.x; r < R; r += kThreads) {
    const double v = sum_pts(D + r * kLD);
    row[b_off + r] = first ? v : row[b_off + r] + v;
  }
}

// The same for a 50 x 50 layer, each of 250 threads forming a 5 x 2 block
// of dW (rows 5 a .. 5 a + 4, columns 2 b, 2 b + 1): 7 row reads per 40
// products instead of 2 per 4.
__device__ __forceinline__ void accum_hidden(const float* __restrict__ D, const float* __restrict__ In, int w_off,
                                             int b_off, double* __restrict__ row, bool first) {
  const int tid = threadIdx.x;
  if (tid < 250) {
    const int r0 = 5 * (tid / 25), c0 = 2 * (tid % 25);
    const int rot = tid & 15;
    float acc[5][2];
#pragma unroll
    for (int a = 0; a < 5; ++a) acc[a][0] = acc[a][1] = 0.f;
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
      const int c = 4 * (q ^ rot);
      const float4 u0 = *reinterpret_cast<const float4*>(In + c0 * kLD + c);
      const float4 u1 = *reinterpret_cast<const float4*>(In + (c0 + 1) * kLD + c);
#pragma unroll
      for (int a = 0; a < 5; ++a) {
        const float4 d = *reinterpret_cast<const float4*>(D + (r0 + a) * kLD + c);
        acc[a][0] += d.x * u0.x + d.y * u0.y + d.z * u0.z + d.w * u0.w;
        acc[a][1] += d.x * u1.x + d.y * u1.y + d.z * u1.z + d.w * u1.w;
      }
    }
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double* dst = row + w_off + (r0 + a) * kH + c0 + b;
        *dst = first ? (double)acc[a][b] : *dst + acc[a][b];
      }
  }
  for (int r = tid; r < kH; r += kThreads) {
    const double v = sum_pts(D + r * kLD);
    row[b_off + r] = first ? v : row[b_off + r] + v;
  }
}

// Deltas of this wave's units one layer down: dr[m] = (sum_r W[r][j] D[r][t])
// for j = 13 w + m, W the wave-blocked copy (row r: 64 floats, this wave's
// 13 units at 16 w), D the tile of the layer's deltas.  Multiplied by
// lrelu'(activation) by the caller.
__device__ __forceinline__ void back_units(const float* __restrict__ Wb, const float* __restrict__ D, int w, int t,
                                           float (&dr)[kUW]) {
#pragma unroll
  for (int m = 0; m < kUW; ++m) dr[m] = 0.f;
  const float* wb = Wb + 16 * w;
#pragma unroll 2
  for (int r = 0; r < kH; ++r) {
    const float d = D[r * kLD + t];
    const float4 a = *reinterpret_cast<const float4*>(wb + kBS * r);
    const float4 b = *reinterpret_cast<const float4*>(wb + kBS * r + 4);
    const float4 c = *reinterpret_cast<const float4*>(wb + kBS * r + 8);
    const float e = wb[kBS * r + 12];
    dr[0] += a.x * d;
    dr[1] += a.y * d;
    dr[2] += a.z * d;
    dr[3] += a.w * d;
    dr[4] += b.x * d;
    dr[5] += b.y * d;
    dr[6] += b.z * d;
    dr[7] += b.w * d;
    dr[8] += c.x * d;
    dr[9] += c.y * d;
    dr[10] += c.z * d;
    dr[11] += c.w * d;
    dr[12] += e * d;
  }
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_ode_bwd(
    const float* __restrict__ y0, int n, const float* __restrict__ params, float4 tau4, float sign, float dt,
    const float* __restrict__ zsave, const float* __restrict__ gy1, float* __restrict__ gy0,
    double* __restrict__ partial) {
  __shared__ __align__(256) float T1[kTile];
  __shared__ __align__(256) float T2[kTile];
  __shared__ __align__(256) float TX[4 * kLD];  // layer-0 input x, transposed
  __shared__ __align__(256) float TD[4 * kLD];  // output deltas, transposed
  __shared__ __align__(16) float P[kBwdP];
  load_params(params, P, true);
  const int w = wave_id(), t = threadIdx.x & (kPts - 1);
  const int j0 = kUW * w, j1 = units_end(w);
  double* row = partial + (size_t)blockIdx.x * kNP;
  const float third = 1.0f / 3.0f;
  const float taus[4] = {tau4.x, tau4.y, tau4.z, tau4.w};
  const int nchunks = (n + kPts - 1) / kPts;
  __syncthreads();
#pragma unroll 1
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const int i = chunk * kPts + t;
    const bool valid = i < n;
    float y[3] = {0.5f, 0.5f, 0.5f}, g[3] = {0.f, 0.f, 0.f}, zs[9];
    for (int c = 0; c < 9; ++c) zs[c] = 0.5f;
    if (valid) {
      for (int c = 0; c < 3; ++c) {
        y[c] = y0[3 * i + c];
        g[c] = gy1[3 * i + c];
      }
      for (int c = 0; c < 9; ++c) zs[c] = zsave[9 * i + c];
    }
    // dL/dk_s from y1 = y0 + (k1 + 3 (k2 + k3) + k4) dt / 8; dL/dy0 from y0
    const float wt = dt * 0.125f;
    float dk[4][3], dy[3];
    for (int c = 0; c < 3; ++c) {
      dk[0][c] = g[c] * wt;
      dk[1][c] = 3.0f * g[c] * wt;
      dk[2][c] = 3.0f * g[c] * wt;
      dk[3][c] = g[c] * wt;
      dy[c] = g[c];
    }
#pragma unroll 1
    for (int s = 3; s >= 0; --s) {
      const bool first = chunk == (int)blockIdx.x && s == 3;
      float z[3], dout[3];
      for (int c = 0; c < 3; ++c) {
        z[c] = s == 0 ? y[c] : zs[3 * (s - 1) + c];
        dout[c] = valid ? sign * (s == 0 ? dk[0][c] : s == 1 ? dk[1][c] : s == 2 ? dk[2][c] : dk[3][c]) : 0.f;
      }
      const float x[4] = {z[0] - 0.5f, z[1] - 0.5f, z[2] - 0.5f, taus[s] - 0.5f};
      if (w == 0) {
        for (int c = 0; c < 4; ++c) TX[c * kLD + t] = x[c];
        for (int c = 0; c < 3; ++c) TD[c * kLD + t] = dout[c];
      }
      float dr[kUW];
      // recompute: T1 <- h1, T2 <- h2, T1 <- h3
      {
        float h[kH];
        layer0_rows(P, x, T1, w, t);
        __syncthreads();
        load_col(T1, t, h);
        hidden_rows(P + kLW1, P + kLB1, h, T2, w, t);
        __syncthreads();
        load_col(T2, t, h);
        hidden_rows(P + kLW2, P + kLB2, h, T1, w, t);
      }
      __syncthreads();
      // output layer: W3 / b3 gradient; delta3 of this wave's units (held until T1 is free)
      accum_layer(TD, T1, 3, kH, kOW3, kOB3, row, first);
#pragma unroll
      for (int m = 0; m < kUW; ++m) {
        const int j = j0 + m;
        if (j < j1)
          dr[m] = (P[kLW3 + j] * dout[0] + P[kLW3 + kRS + j] * dout[1] + P[kLW3 + 2 * kRS + j] * dout[2]) *
                  dlrelu(T1[j * kLD + t]);
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kUW; ++m)
        if (j0 + m < j1) T1[(j0 + m) * kLD + t] = dr[m];   // T1 <- delta3
      __syncthreads();
      // layer 2: gradient from (delta3, h2); delta2 of this wave's units
      accum_hidden(T1, T2, kOW2, kOB2, row, first);
      back_units(P + kLW2B, T1, w, t, dr);
#pragma unroll
      for (int m = 0; m < kUW; ++m)
        if (j0 + m < j1) dr[m] *= dlrelu(T2[(j0 + m) * kLD + t]);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kUW; ++m)
        if (j0 + m < j1) T2[(j0 + m) * kLD + t] = dr[m];   // T2 <- delta2
      layer0_rows(P, x, T1, w, t);                         // T1 <- h1 again
      __syncthreads();
      // layer 1: gradient from (delta2, h1); delta1 of this wave's units
      accum_hidden(T2, T1, kOW1, kOB1, row, first);
      back_units(P + kLW1B, T2, w, t, dr);
#pragma unroll
      for (int m = 0; m < kUW; ++m)
        if (j0 + m < j1) dr[m] *= dlrelu(T1[(j0 + m) * kLD + t]);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kUW; ++m)
        if (j0 + m < j1) T1[(j0 + m) * kLD + t] = dr[m];   // T1 <- delta1
      __syncthreads();
      // layer 0: gradient from (delta1, x); dL/dz = (W0^T delta1)[0:3]
      accum_layer(T1, TX, kH, 4, kOW0, kOB0, row, first);
      float dz[3] = {0.f, 0.f, 0.f};
#pragma unroll 5
      for (int r = 0; r < kH; ++r) {
        const float d = T1[r * kLD + t];
        const float4 u = *reinterpret_cast<const float4*>(P + kLW0 + 4 * r);
        dz[0] += u.x * d;
        dz[1] += u.y * d;
        dz[2] += u.z * d;
      }
      __syncthreads();
      // chain into the earlier stages and y0
      for (int c = 0; c < 3; ++c) {
        dy[c] += dz[c];
        if (s == 3) {          // z4 = y0 + dt (k1 - k2 + k3)
          dk[0][c] += dt * dz[c];
          dk[1][c] -= dt * dz[c];
          dk[2][c] += dt * dz[c];
        } else if (s == 2) {   // z3 = y0 + dt (k2 - k1 / 3)
          dk[1][c] += dt * dz[c];
          dk[0][c] -= dt * third * dz[c];
        } else if (s == 1) {   // z2 = y0 + dt k1 / 3
          dk[0][c] += dt * third * dz[c];
        }
      }
    }
    if (valid && gy0 && w == 0)
      for (int c = 0; c < 3; ++c) gy0[3 * i + c] = dy[c];
  }
}

// out[e] = sum over blocks of partial[b][e] (FP64).  A block covers 64
// elements; wave w adds the rows b = w, w + 8, ... in order, and the eight
// wave sums are added in wave order: a fixed order for any scheduling.
constexpr int kRedWaves = 8;
__global__ __launch_bounds__(64 * kRedWaves) void k_ode_reduce(const double* __restrict__ partial, int nblocks,
                                                              float* __restrict__ out) {
  __shared__ double S[kRedWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (e < kNP) {
#pragma unroll 4
    for (int b = w; b < nblocks; b += kRedWaves) s += partial[(size_t)b * kNP + e];
  }
  S[w][lane] = s;
  __syncthreads();
  if (w == 0 && e < kNP) {
    double t = S[0][lane];
    for (int k = 1; k < kRedWaves; ++k) t += S[k][lane];
    out[e] = (float)t;
  }
}

int bwd_blocks(int64_t n) { return (int)std::min<int64_t>((n + kPts - 1) / kPts, kMaxBlocks); }

}  // namespace
}  // namespace mo

using namespace mo;

MO_EXPORT int64_t mo_ode_rk4_workspace_bytes(int64_t n) {
  return n <= 0 ? 0 : (int64_t)bwd_blocks(n) * kNP * (int64_t)sizeof(double);
}

MO_EXPORT int mo_ode_rk4_forward(const float* y0, int64_t n, const float* params, const float* tau4, float sign,
                                 float dt, float* y1, float* zsave, void* stream) {
  MO_CHECK(params && tau4 && n >= 0 && n <= (1LL << 30) && (n == 0 || (y0 && y1)), MO_ERR_ARG,
           "mo_ode_rk4_forward: bad arguments");
  if (n == 0) return MO_OK;
  const float4 tau = make_float4(tau4[0], tau4[1], tau4[2], tau4[3]);
  hipLaunchKernelGGL(k_ode_fwd, dim3((unsigned)((n + kPts - 1) / kPts)), dim3(kThreads), 0, (hipStream_t)stream, y0,
                     (int)n, params, tau, sign, dt, y1, zsave);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int mo_ode_rk4_backward(const float* y0, int64_t n, const float* params, const float* tau4, float sign,
                                  float dt, const float* zsave, const float* gy1, float* gy0, float* gparams,
                                  void* work, int64_t work_bytes, void* stream) {
  MO_CHECK(params && tau4 && gparams && n >= 0 && n <= (1LL << 30) && (n == 0 || (y0 && zsave && gy1)), MO_ERR_ARG,
           "mo_ode_rk4_backward: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    MO_HIP(hipMemsetAsync(gparams, 0, kNP * sizeof(float), s));
    return MO_OK;
  }
  const int nb = bwd_blocks(n);
  const int64_t need = (int64_t)nb * kNP * (int64_t)sizeof(double);
  MO_CHECK(!work || work_bytes >= need, MO_ERR_ARG, "mo_ode_rk4_backward: workspace of %lld bytes, %lld needed",
           (long long)work_bytes, (long long)need);
  MO_CHECK(((uintptr_t)work & 7) == 0, MO_ERR_ARG, "mo_ode_rk4_backward: workspace not 8-byte aligned");
  double* partial = (double*)work;
  if (!partial) {
    const hipError_t e = hipMallocAsync((void**)&partial, (size_t)need, s);
    MO_CHECK(e == hipSuccess, MO_ERR_NOMEM, "mo_ode_rk4_backward: hipMallocAsync failed: %s", hipGetErrorString(e));
  }
  const float4 tau = make_float4(tau4[0], tau4[1], tau4[2], tau4[3]);
  hipLaunchKernelGGL(k_ode_bwd, dim3(nb), dim3(kThreads), 0, s, y0, (int)n, params, tau, sign, dt, zsave, gy1, gy0,
                     partial);
  hipLaunchKernelGGL(k_ode_reduce, dim3((kNP + 63) / 64), dim3(64 * kRedWaves), 0, s, partial, nb, gparams);
  const hipError_t e = hipGetLastError();
  if (!work) (void)hipFreeAsync(partial, s);
  MO_CHECK(e == hipSuccess, MO_ERR_HIP, "mo_ode_rk4_backward: launch failed: %s", hipGetErrorString(e));
  return MO_OK;
}
