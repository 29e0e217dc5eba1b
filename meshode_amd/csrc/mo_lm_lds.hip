// Register/LDS-resident variant of the batched LM solver (same algorithm as
// mo_lm.hip: Deformer::Deform / DeformGraph, deformer.cc:18-92, 370-442,
// with Ceres' trust-region LM restated and a matrix-free block-Jacobi PCG).
//
// One workgroup of T threads solves one pair whose nv <= T * VPT vertices
// are statically owned (vertex i by thread i % T, slot i / T).  Everything
// the PCG iterates on lives in VGPRs for the whole solve: the distance
// gradient g, the residual r, the direction p, q = A p, the step delta and
// d = lambda^2 deg + LM diagonal / radius (the diagonal of A apart from
// g g^T, which is also the block-Jacobi preconditioner's diagonal).  The
// one vector neighbours read (u while evaluating the Jacobian, z = M^-1 r
// inside PCG, u + delta for the candidate) is published to a
// float4-per-vertex LDS array, and the pair's adjacency is copied once into
// LDS as uint16 entries (neighbour << 1 | multiplicity - 1), so the PCG loop
// touches no global memory.  Reductions are FP64, through DPP lane moves
// and one LDS exchange between waves.
//
// Differences from the global kernel that do not change the algorithm:
//   * the preconditioner M = g g^T + diag(d) is applied with the
//     Sherman-Morrison formula each iteration instead of a stored inverse;
//   * A p is formed as A z + beta A p_prev (p = z + beta p_prev), so only z
//     is published and an iteration needs two barriers;
//   * the rigidity residuals are linear in x, so the rigidity part of the
//     model cost change equals its actual change E_r(u + delta) - E_r(u);
//     only u + delta has to be gathered for the candidate.
#include "mo_common.h"
#include "mo_lm.h"
#include "mo_lm_util.h"
#include "mo_lookup.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace mo {

namespace {

enum GSlot { gGR = 1, gS = 2, gDG = 3 };   // global slots of LMArrays::work used here

// Block-wide exclusive scan of one int per thread (T <= 1024).
__device__ inline int block_exclusive_scan(int x, int* buf, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int v = x;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) buf[w] = v;
  __syncthreads();
  int base = 0;
  total = 0;
  for (int q = 0; q < nw; ++q) {
    if (q < w) base += buf[q];
    total += buf[q];
  }
  __syncthreads();
  return base + v - x;
}

// An index the optimizer cannot see through: per-vertex addresses are then
// formed at each use (base + index) instead of being hoisted out of the
// solver loops, where they would hold dozens of registers for the whole
// kernel.
__device__ inline int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <int T, int VPT>
__global__ void __launch_bounds__(T) k_lm_solve_lds(const LMPairDesc* __restrict__ pairs,
                                                    const int* __restrict__ plist, LMArrays A, LMOptionsDev opt,
                                                    LMResultDev* __restrict__ res, long long totalV) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[2 * 8 * kMaxWaves];
  __shared__ int scan_buf[kMaxWaves];
  const double o_min_diag = opt.min_lm_diagonal, o_max_diag = opt.max_lm_diagonal;
  const float o_pcg_tol = opt.pcg_relative_tolerance;
  const int o_pcg_max = opt.pcg_max_iterations;
  const int pid = plist[blockIdx.x];
  const LMPairDesc P = pairs[pid];
  const int tid = threadIdx.x;
  const int nv = P.nv;
  const long long vb = P.vbase;
  const float l2 = P.lambda * P.lambda;
  const float* __restrict__ grid = P.grid;
  const int N = P.N;
  const float4* __restrict__ X0 = A.x0 + vb;
  float4* __restrict__ U = A.u + vb;
  const int2* __restrict__ ADJ = A.adj + vb;
  const unsigned* __restrict__ NB = A.nbr + P.ebase;
  float4* __restrict__ GR = A.work + (size_t)gGR * totalV + vb;
  float4* __restrict__ S = A.work + (size_t)gS * totalV + vb;
  float4* __restrict__ DG = A.work + (size_t)gDG * totalV + vb;
  float4* VEC = (float4*)smem;
  unsigned short* LCSR = (unsigned short*)(smem + (size_t)nv * sizeof(float4));
  int sel = 0;
  const long long clock_start = wall_clock64();

  // ---- adjacency -> LDS (packed, compacted by a block scan) -------------
  unsigned csr[VPT];   // (LDS offset << 16) | count
  {
    int cnt[VPT], gst[VPT], mine = 0;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      cnt[k] = 0;
      gst[k] = 0;
      if (i < nv) {
        const int2 a = ADJ[i];
        gst[k] = a.x;
        cnt[k] = a.y;
        mine += a.y;
      }
    }
    int total;
    int off = block_exclusive_scan(mine, scan_buf, total);
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      csr[k] = ((unsigned)off << 16) | (unsigned)cnt[k];
      for (int e = 0; e < cnt[k]; ++e) {
        const unsigned en = NB[gst[k] + e];
        LCSR[off + e] = (unsigned short)((nb_index(en) << 1) | ((en >> 24) - 1u));
      }
      off += cnt[k];
    }
  }

  // ---- per-vertex register state ----------------------------------------
  // g: distance gradient; r: PCG residual; p: direction; q = A p;
  // d = lambda^2 deg + LM diagonal / radius (the diagonal of A without
  // g g^T, also the preconditioner's diagonal); e: the step delta.
  float gx[VPT], gy[VPT], gz[VPT];
  float rx[VPT], ry[VPT], rz_[VPT];
  float px[VPT], py[VPT], pz[VPT];
  float qx[VPT], qy[VPT], qz[VPT];
  float dx[VPT], dy[VPT], dz[VPT];
  float ex[VPT], ey[VPT], ez[VPT];

  // z = M^-1 r with M = g g^T + diag(d) (Sherman-Morrison)
  auto precond = [&](int k, float r0, float r1, float r2, float& z0, float& z1, float& z2) {
    const float i0 = __builtin_amdgcn_rcpf(dx[k]), i1 = __builtin_amdgcn_rcpf(dy[k]),
                i2 = __builtin_amdgcn_rcpf(dz[k]);
    const float h0 = gx[k] * i0, h1 = gy[k] * i1, h2 = gz[k] * i2;
    const float sig = 1.0f + gx[k] * h0 + gy[k] * h1 + gz[k] * h2;
    const float c = (h0 * r0 + h1 * r1 + h2 * r2) * __builtin_amdgcn_rcpf(sig);
    z0 = r0 * i0 - h0 * c;
    z1 = r1 * i1 - h1 * c;
    z2 = r2 * i2 - h2 * c;
  };

  // sum over the vertex's neighbours of multiplicity * VEC[neighbour].xyz
  auto gather = [&](int k, float& sx, float& sy, float& sz) {
    sx = 0.f; sy = 0.f; sz = 0.f;
    const int cs = opaque((int)csr[k]);
    const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
    #pragma unroll 1
    for (int q = 0; q < c; ++q) {
      const unsigned w = LCSR[o + q];
      const float m = (float)((w & 1u) + 1u);
      const float4 vj = VEC[w >> 1];
      sx += m * vj.x; sy += m * vj.y; sz += m * vj.z;
    }
  };

  auto eval = [&](bool first, double& cv, double& ce, double& gmax) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      if (i < nv) VEC[i] = U[i];
    }
    __syncthreads();
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      if (i >= nv) continue;
      const float4 x0 = X0[i];
      const float4 ui = VEC[i];
      const JetF J = lookup_value_grad(grid, N, x0.x + ui.x, x0.y + ui.y, x0.z + ui.z);
      gx[k] = J.v0; gy[k] = J.v1; gz[k] = J.v2;
      const float ldk = l2 * x0.w;
      float sx = 0.f, sy = 0.f, sz = 0.f;
      double e = 0.0;
      const int cs = opaque((int)csr[k]);
      const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
      #pragma unroll 1
      for (int q = 0; q < c; ++q) {
        const unsigned w = LCSR[o + q];
        const float m = (float)((w & 1u) + 1u);
        const float4 uj = VEC[w >> 1];
        const float ax = ui.x - uj.x, ay = ui.y - uj.y, az = ui.z - uj.z;
        sx += m * ax; sy += m * ay; sz += m * az;
        e += (double)m * ((double)ax * ax + (double)ay * ay + (double)az * az);
      }
      const float g0 = J.a * J.v0 + l2 * sx, g1 = J.a * J.v1 + l2 * sy, g2 = J.a * J.v2 + l2 * sz;
      GR[i] = make_float4(g0, g1, g2, J.a);   // J^T f and the residual d
      v[0] += 0.5 * (double)J.a * (double)J.a;
      v[1] += 0.25 * (double)l2 * e;
      double gm = fabs((double)g0);
      gm = fabs((double)g1) > gm ? fabs((double)g1) : gm;
      gm = fabs((double)g2) > gm ? fabs((double)g2) : gm;
      v[2] = gm > v[2] ? gm : v[2];
      if (first) {
        const float c0 = J.v0 * J.v0 + ldk, c1 = J.v1 * J.v1 + ldk, c2 = J.v2 * J.v2 + ldk;
        S[i] = make_float4(1.0f / (1.0f + sqrtf(c0)), 1.0f / (1.0f + sqrtf(c1)), 1.0f / (1.0f + sqrtf(c2)), 0.f);
      }
    }
    block_reduce_dpp<2, 1>(v, red, sel);
    cv = v[0]; ce = v[1]; gmax = v[2];
  };

  // PCG on (J^T J + D^2 / radius) delta = -J^T f.  Each iteration publishes
  // z = M^-1 r once and forms A p as A z + beta A p_prev (p = z + beta
  // p_prev), so an iteration has two barriers: the reduction of p.Ap, and
  // the reduction of r.z and r.r, which also publishes the new z.
  auto solve = [&](bool reuse, double radius) -> int {
    const float finv_radius = (float)(1.0 / radius);
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      px[k] = py[k] = pz[k] = 0.f;
      qx[k] = qy[k] = qz[k] = 0.f;
      ex[k] = ey[k] = ez[k] = 0.f;
      rx[k] = ry[k] = rz_[k] = 0.f;
      dx[k] = dy[k] = dz[k] = 1.f;
      if (i >= nv) continue;
      const float4 s = S[i];
      const float ldk = l2 * X0[i].w;
      float4 dg;
      if (!reuse) {
        const double cc[3] = {(double)gx[k] * gx[k] + ldk, (double)gy[k] * gy[k] + ldk,
                              (double)gz[k] * gz[k] + ldk};
        const double sc[3] = {s.x, s.y, s.z};
        float d3[3];
        for (int a = 0; a < 3; ++a) {
          double t = sc[a] * sc[a] * cc[a];
          t = t < o_min_diag ? o_min_diag : (t > o_max_diag ? o_max_diag : t);
          d3[a] = (float)t;
        }
        dg = make_float4(d3[0], d3[1], d3[2], 0.f);
        DG[i] = dg;
      } else {
        dg = DG[i];
      }
      // LM regularization D^2 / radius in the unscaled variables (the
      // diagonal over s^2), capped at 1e30; FP32 with a reciprocal (the
      // FP64 divisions' temporaries would set the kernel's register peak)
      dx[k] = ldk + fminf(dg.x * __builtin_amdgcn_rcpf(s.x * s.x) * finv_radius, 1e30f);
      dy[k] = ldk + fminf(dg.y * __builtin_amdgcn_rcpf(s.y * s.y) * finv_radius, 1e30f);
      dz[k] = ldk + fminf(dg.z * __builtin_amdgcn_rcpf(s.z * s.z) * finv_radius, 1e30f);
      const float4 gr = GR[i];
      rx[k] = -gr.x; ry[k] = -gr.y; rz_[k] = -gr.z;
      float z0, z1, z2;
      precond(k, rx[k], ry[k], rz_[k], z0, z1, z2);
      VEC[i] = make_float4(z0, z1, z2, 0.f);
      v[0] += (double)(rx[k] * z0 + ry[k] * z1 + rz_[k] * z2);
      v[1] += (double)(rx[k] * rx[k] + ry[k] * ry[k] + rz_[k] * rz_[k]);
    }
    block_reduce_dpp<2, 0>(v, red, sel);   // also publishes z
    double rz = v[0];
    const double rr0 = v[1];
    if (!(rr0 > 0.0)) return 0;
    const double tol2 = (double)o_pcg_tol * o_pcg_tol * rr0;
    float beta = 0.f;
    int it = 0;
    while (it < o_pcg_max) {
      double w1[1] = {0.0};
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int i = opaque(tid + k * T);
        if (i >= nv) continue;
        const float4 zi = VEC[i];
        float sx, sy, sz;
        gather(k, sx, sy, sz);
        const float gzz = gx[k] * zi.x + gy[k] * zi.y + gz[k] * zi.z;
        const float wx = gx[k] * gzz + dx[k] * zi.x - l2 * sx;
        const float wy = gy[k] * gzz + dy[k] * zi.y - l2 * sy;
        const float wz = gz[k] * gzz + dz[k] * zi.z - l2 * sz;
        px[k] = zi.x + beta * px[k]; py[k] = zi.y + beta * py[k]; pz[k] = zi.z + beta * pz[k];
        qx[k] = wx + beta * qx[k]; qy[k] = wy + beta * qy[k]; qz[k] = wz + beta * qz[k];
        w1[0] += (double)(px[k] * qx[k] + py[k] * qy[k] + pz[k] * qz[k]);
      }
      block_reduce_dpp<1, 0>(w1, red, sel);
      const double pq = w1[0];
      if (!(pq > 0.0)) break;
      const float fa = (float)(rz / pq);
      double w2[2] = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int i = opaque(tid + k * T);
        if (i >= nv) continue;
        ex[k] += fa * px[k]; ey[k] += fa * py[k]; ez[k] += fa * pz[k];
        rx[k] -= fa * qx[k]; ry[k] -= fa * qy[k]; rz_[k] -= fa * qz[k];
        float z0, z1, z2;
        precond(k, rx[k], ry[k], rz_[k], z0, z1, z2);
        VEC[i] = make_float4(z0, z1, z2, 0.f);
        w2[0] += (double)(rx[k] * z0 + ry[k] * z1 + rz_[k] * z2);
        w2[1] += (double)(rx[k] * rx[k] + ry[k] * ry[k] + rz_[k] * rz_[k]);
      }
      block_reduce_dpp<2, 0>(w2, red, sel);   // also publishes z
      ++it;
      if (w2[1] <= tol2) break;
      beta = (float)(w2[0] / rz);
      rz = w2[0];
    }
    return it;
  };

  double ce_now = 0.0;   // rigidity energy at the current x (set by eval)
  auto candidate = [&]() -> LMCandidate {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      if (i < nv) {
        const float4 ui = U[i];
        VEC[i] = make_float4(ui.x + ex[k], ui.y + ey[k], ui.z + ez[k], 0.f);
      }
    }
    __syncthreads();
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      if (i >= nv) continue;
      const float4 x0 = X0[i];
      const float dres = GR[i].w;
      const float gd = gx[k] * ex[k] + gy[k] * ey[k] + gz[k] * ez[k];
      v[0] += (double)gd * ((double)dres + 0.5 * (double)gd);
      const float4 wv = VEC[i];
      const float wx = wv.x, wy = wv.y, wz = wv.z;
      const float dc = lookup_f32(grid, N, x0.x + wx, x0.y + wy, x0.z + wz);
      v[1] += 0.5 * (double)dc * (double)dc;
      double e = 0.0;
      const int cs = opaque((int)csr[k]);
      const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
      #pragma unroll 1
      for (int q = 0; q < c; ++q) {
        const unsigned w = LCSR[o + q];
        const float m = (float)((w & 1u) + 1u);
        const float4 wj = VEC[w >> 1];
        const float ax = wx - wj.x, ay = wy - wj.y, az = wz - wj.z;
        e += (double)m * ((double)ax * ax + (double)ay * ay + (double)az * az);
      }
      v[2] += 0.25 * (double)l2 * e;
      v[3] += (double)ex[k] * ex[k] + (double)ey[k] * ey[k] + (double)ez[k] * ez[k];
      const float4 ui = U[i];
      const float qx0 = x0.x + ui.x, qy0 = x0.y + ui.y, qz0 = x0.z + ui.z;
      v[4] += (double)qx0 * qx0 + (double)qy0 * qy0 + (double)qz0 * qz0;
    }
    block_reduce_dpp<5, 0>(v, red, sel);
    // rigidity residuals are linear: their model change equals the actual one
    return LMCandidate{-v[0] - (v[2] - ce_now), v[1] + v[2], sqrt(v[3]), sqrt(v[4])};
  };

  auto eval_tracked = [&](bool first, double& cv, double& ce, double& gmax) {
    eval(first, cv, ce, gmax);
    ce_now = ce;
  };
  auto accept = [&]() {
    // u <- u + delta: exactly the candidate published to LDS
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int i = opaque(tid + k * T);
      if (i < nv) U[i] = VEC[i];
    }
  };
  lm_trust_region_loop(opt, eval_tracked, solve, candidate, accept, res + pid, clock_start);
}

template <int T, int VPT>
int launch_bucket(const LMPairDesc* pairs, const int* plist, int count, const LMArrays& A, const LMOptionsDev& opt,
                  LMResultDev* res, long long totalV, size_t lds_bytes, hipStream_t s) {
  auto kern = k_lm_solve_lds<T, VPT>;
  MO_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  hipLaunchKernelGGL(kern, dim3(count), dim3(T), lds_bytes, s, pairs, plist, A, opt, res, totalV);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

}  // namespace

// Shapes of the register/LDS kernel: (threads, vertices per thread).  A pair
// of nv vertices can use any shape with T * VPT >= nv.  Which one it gets:
//   1. the size-class table, when one is set (lm_lds_set_table: pairs with
//      nv <= max_nv[r], first r in ascending order, use shape[r]; this is
//      what the autotuner installs after timing the candidates);
//   2. else the first fitting shape in the preference order: the built-in
//      order below, or MO_LM_SHAPES="TxVPT,TxVPT,..." read once.
static const int kShapeT[] = {1024, 1024, 1024, 1024, 512, 512, 512, 512, 256, 256, 256};
static const int kShapeVPT[] = {1, 2, 4, 5, 8, 10, 12, 16, 20, 24, 32};
constexpr int kNumShapes = 11;

static std::mutex g_shape_mu;
static std::vector<int> g_table_nv, g_table_shape;   // guarded by g_shape_mu

static const std::vector<int>& shape_preference() {
  static const std::vector<int> pref = [] {
    std::vector<int> p;
    const char* env = getenv("MO_LM_SHAPES");
    if (env) {
      std::string e(env);
      size_t pos = 0;
      while (pos < e.size()) {
        size_t comma = e.find(',', pos);
        std::string tok = e.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
        int t = 0, v = 0;
        if (sscanf(tok.c_str(), "%dx%d", &t, &v) == 2)
          for (int b = 0; b < kNumShapes; ++b)
            if (kShapeT[b] == t && kShapeVPT[b] == v) p.push_back(b);
        if (comma == std::string::npos) break;
        pos = comma + 1;
      }
    }
    if (p.empty())
      for (int b = 0; b < kNumShapes; ++b) p.push_back(b);
    return p;
  }();
  return pref;
}

int lm_lds_buckets() { return kNumShapes; }

int lm_lds_shape(int b, int* threads, int* vpt) {
  if (b < 0 || b >= kNumShapes) return fail(MO_ERR_ARG, "shape index %d out of range [0, %d)", b, kNumShapes);
  if (threads) *threads = kShapeT[b];
  if (vpt) *vpt = kShapeVPT[b];
  return MO_OK;
}

int lm_lds_set_table(const int* max_nv, const int* shape, int n) {
  MO_CHECK(n >= 0 && (n == 0 || (max_nv && shape)), MO_ERR_ARG, "shape table: bad arguments");
  for (int r = 0; r < n; ++r) {
    MO_CHECK(shape[r] >= 0 && shape[r] < kNumShapes, MO_ERR_ARG, "shape table row %d: shape %d out of range", r,
             shape[r]);
    MO_CHECK(max_nv[r] >= 1 && (r == 0 || max_nv[r] > max_nv[r - 1]), MO_ERR_ARG,
             "shape table row %d: max_nv must be positive and ascending", r);
    MO_CHECK((long long)kShapeT[shape[r]] * kShapeVPT[shape[r]] >= max_nv[r], MO_ERR_ARG,
             "shape table row %d: shape %dx%d holds fewer than %d vertices", r, kShapeT[shape[r]],
             kShapeVPT[shape[r]], max_nv[r]);
  }
  std::lock_guard<std::mutex> lk(g_shape_mu);
  g_table_nv.assign(max_nv, max_nv + n);
  g_table_shape.assign(shape, shape + n);
  return MO_OK;
}

int lm_lds_bucket_for(int nv) {
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    for (size_t r = 0; r < g_table_nv.size(); ++r)
      if (nv <= g_table_nv[r]) return g_table_shape[r];
  }
  for (int b : shape_preference())
    if (nv <= kShapeT[b] * kShapeVPT[b]) return b;
  return -1;
}

size_t lm_lds_bytes(int nv, long long entries) {
  return (((size_t)nv * 16 + (size_t)entries * 2) + 15) & ~(size_t)15;
}

int lm_lds_launch(int bucket, const LMPairDesc* pairs, const int* plist, int count, const LMArrays& A,
                  const LMOptionsDev& opt, LMResultDev* res, long long totalV, size_t lds_bytes, hipStream_t s) {
  if (count <= 0) return MO_OK;
  switch (bucket) {
#define MO_SHAPE(ID, T, V) \
    case ID: return launch_bucket<T, V>(pairs, plist, count, A, opt, res, totalV, lds_bytes, s);
    MO_SHAPE(0, 1024, 1) MO_SHAPE(1, 1024, 2) MO_SHAPE(2, 1024, 4) MO_SHAPE(3, 1024, 5)
    MO_SHAPE(4, 512, 8) MO_SHAPE(5, 512, 10) MO_SHAPE(6, 512, 12) MO_SHAPE(7, 512, 16)
    MO_SHAPE(8, 256, 20) MO_SHAPE(9, 256, 24) MO_SHAPE(10, 256, 32)
#undef MO_SHAPE
    default: return fail(MO_ERR_ARG, "bad LDS solver shape %d", bucket);
  }
}

}  // namespace mo
