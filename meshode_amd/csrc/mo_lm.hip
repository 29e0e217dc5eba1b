// Batched Levenberg-Marquardt solve of Deformer::Deform / DeformGraph
// (deformer.cc:18-92, 370-442) on gfx950: one workgroup per pair, the whole
// trust-region loop inside one persistent kernel.
//
// Problem (unchanged from the reference): per vertex u the DistanceLoss
// residual d(x_u) (distanceloss.h:6-25), per directed edge (a, b) the
// EdgeLoss residual lambda (x_a - x_b - v_ab) with v_ab = x0_a - x0_b
// (edgeloss.h:8-33, deformer.cc:44).  Unknowns are kept as displacements
// u = x - x0, so every edge residual is exactly lambda (u_a - u_b) and the
// rigidity energy of all directed edges between u and w collapses to
// 1/2 lambda^2 m_uw ||u_u - u_w||^2 with m_uw the number of such edges.  The
// adjacency is stored once per vertex (neighbour, multiplicity).
//
// Minimizer: Ceres' trust-region loop with the Levenberg-Marquardt strategy
// (Jacobi scaling fixed at iteration 0, LM diagonal clamp(colnorm^2(J s)) /
// radius, rho test against min_relative_decrease, radius update
// max(1/3, 1 - (2 rho - 1)^3), parameter / function / gradient tolerances,
// 5 consecutive invalid steps = failure).  Ceres' exact sparse Cholesky
// (SPARSE_NORMAL_CHOLESKY, deformer.cc:57) is replaced by a matrix-free
// preconditioned conjugate gradient on
//     (J^T J + D^2) delta = -J^T f,
//     J^T J = blockdiag(g g^T) + lambda^2 (L_m (x) I_3)
// in the unscaled variables (Jacobi scaling cancels under the 3x3
// block-Jacobi preconditioner), with FP32 vectors and FP64 reductions.
//
// All scalar state of the trust region is replicated in every thread and
// updated from block-wide reductions that every thread sums in the same
// order, so control flow is uniform without any broadcast.
#include "mo_common.h"
#include "mo_lm.h"
#include "mo_lm_util.h"
#include "mo_lookup.h"

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <mutex>

namespace mo {

namespace {

// Work slots (float4 each, SoA over all vertices of the batch).
enum Slot { kG = 0, kGR, kS, kDG, kDT, kMI0, kMI1, kDL, kR, kZ, kP0, kP1, kQ, kNumSlots };
static_assert(kNumSlots == kLMWorkSlots, "work slot count is part of the host contract");

template <class Team>
__device__ inline void lm_solve_pair(const LMPairDesc P, LMResultDev* out, const LMArrays& A,
                                     const LMOptionsDev& opt, long long totalV, Team& team, double* red) {
  const int T = blockDim.x, tid = threadIdx.x;
  const int i0 = team.rank() * T + tid, stride = team.size() * T;
  const int nv = P.nv;
  const long long vb = P.vbase;
  const float l2 = P.lambda * P.lambda;
  const bool weighted = (P.flags & 1) != 0;
  const float* __restrict__ WT = A.wt + P.ebase;
  const float* __restrict__ grid = P.grid;
  const int N = P.N;
  const float4* __restrict__ X0 = A.x0 + vb;
  float4* __restrict__ U = A.u + vb;
  const int2* __restrict__ ADJ = A.adj + vb;
  const unsigned* __restrict__ NB = A.nbr + P.ebase;
  float4* __restrict__ W[kNumSlots];
  for (int k = 0; k < kNumSlots; ++k) W[k] = A.work + (size_t)k * totalV + vb;
  float4* __restrict__ G = W[kG];
  float4* __restrict__ GR = W[kGR];
  float4* __restrict__ S = W[kS];
  float4* __restrict__ DG = W[kDG];
  float4* __restrict__ DT = W[kDT];
  float4* __restrict__ MI0 = W[kMI0];
  float4* __restrict__ MI1 = W[kMI1];
  float4* __restrict__ DL = W[kDL];
  float4* __restrict__ R = W[kR];
  float4* __restrict__ Z = W[kZ];
  float4* __restrict__ PB[2] = {W[kP0], W[kP1]};
  float4* __restrict__ Q = W[kQ];
  int sel = 0;
  const long long clock_start = wall_clock64();

  // ---- evaluate residuals, gradient J^T f, cost at the current u --------
  auto eval_jacobian = [&](bool first, double& cv, double& ce, double& gmax) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < nv; i += stride) {
      const float4 x0 = X0[i];
      const float4 ui = U[i];
      const JetF J = lookup_jet(grid, N, x0.x + ui.x, x0.y + ui.y, x0.z + ui.z);
      G[i] = make_float4(J.v0, J.v1, J.v2, J.a);
      float rx = 0.f, ry = 0.f, rz = 0.f;
      double e = 0.0;
      const int2 a = ADJ[i];
      for (int k = 0; k < a.y; ++k) {
        const unsigned en = NB[a.x + k];
        const float m = weighted ? WT[a.x + k] : nb_mult(en);
        const float4 uj = U[nb_index(en)];
        const float dx = ui.x - uj.x, dy = ui.y - uj.y, dz = ui.z - uj.z;
        rx += m * dx; ry += m * dy; rz += m * dz;
        e += (double)m * ((double)dx * dx + (double)dy * dy + (double)dz * dz);
      }
      const float gx = J.a * J.v0 + l2 * rx, gy = J.a * J.v1 + l2 * ry, gz = J.a * J.v2 + l2 * rz;
      GR[i] = make_float4(gx, gy, gz, 0.f);
      v[0] += 0.5 * (double)J.a * (double)J.a;
      v[1] += 0.25 * (double)l2 * e;
      double gm = fabs((double)gx);
      gm = fabs((double)gy) > gm ? fabs((double)gy) : gm;
      gm = fabs((double)gz) > gm ? fabs((double)gz) : gm;
      v[2] = gm > v[2] ? gm : v[2];
      if (first) {
        const float deg = x0.w;
        const float c0 = J.v0 * J.v0 + l2 * deg, c1 = J.v1 * J.v1 + l2 * deg, c2 = J.v2 * J.v2 + l2 * deg;
        S[i] = make_float4(1.0f / (1.0f + sqrtf(c0)), 1.0f / (1.0f + sqrtf(c1)), 1.0f / (1.0f + sqrtf(c2)), 0.f);
      }
    }
    team.template reduce<2, 1>(v, red, sel);
    cv = v[0]; ce = v[1]; gmax = v[2];
  };

  // ---- LM diagonal, block-Jacobi preconditioner, PCG start -------------
  auto setup_linear = [&](bool reuse, double radius, double& rz_out, double& rr_out) {
    double v[2] = {0.0, 0.0};
    const double inv_radius = 1.0 / radius;
    for (int i = i0; i < nv; i += stride) {
      const float4 g = G[i];
      const float deg = X0[i].w;
      const float4 s = S[i];
      const float ld = l2 * deg;
      float4 dg;
      if (!reuse) {
        const double c[3] = {(double)g.x * g.x + ld, (double)g.y * g.y + ld, (double)g.z * g.z + ld};
        const double sc[3] = {s.x, s.y, s.z};
        float d3[3];
        for (int k = 0; k < 3; ++k) {
          double t = sc[k] * sc[k] * c[k];
          t = t < opt.min_lm_diagonal ? opt.min_lm_diagonal : (t > opt.max_lm_diagonal ? opt.max_lm_diagonal : t);
          d3[k] = (float)t;
        }
        dg = make_float4(d3[0], d3[1], d3[2], 0.f);
        DG[i] = dg;
      } else {
        dg = DG[i];
      }
      // D^2 in unscaled variables: diag / (radius s^2)
      double t0 = (double)dg.x / ((double)s.x * s.x) * inv_radius;
      double t1 = (double)dg.y / ((double)s.y * s.y) * inv_radius;
      double t2 = (double)dg.z / ((double)s.z * s.z) * inv_radius;
      const double cap = 1e30;
      const float4 dt = make_float4((float)(t0 < cap ? t0 : cap), (float)(t1 < cap ? t1 : cap),
                                    (float)(t2 < cap ? t2 : cap), 0.f);
      DT[i] = dt;
      // M = g g^T + lambda^2 deg I + D^2, inverted in FP64
      const double a = (double)g.x * g.x + ld + dt.x, b = (double)g.x * g.y, c = (double)g.x * g.z;
      const double d = (double)g.y * g.y + ld + dt.y, e = (double)g.y * g.z, f = (double)g.z * g.z + ld + dt.z;
      const double A0 = d * f - e * e, B0 = c * e - b * f, C0 = b * e - c * d;
      const double det = a * A0 + b * B0 + c * C0;
      float4 m0, m1;
      if (det > 0.0 && det < INFINITY) {
        const double id = 1.0 / det;
        m0 = make_float4((float)(A0 * id), (float)(B0 * id), (float)(C0 * id), (float)((a * f - c * c) * id));
        m1 = make_float4((float)((b * c - a * e) * id), (float)((a * d - b * b) * id), 0.f, 0.f);
      } else {
        m0 = make_float4(a > 0 ? (float)(1.0 / a) : 0.f, 0.f, 0.f, d > 0 ? (float)(1.0 / d) : 0.f);
        m1 = make_float4(0.f, f > 0 ? (float)(1.0 / f) : 0.f, 0.f, 0.f);
      }
      MI0[i] = m0;
      MI1[i] = m1;
      const float4 gr = GR[i];
      const float r0 = -gr.x, r1 = -gr.y, r2 = -gr.z;
      const float z0 = m0.x * r0 + m0.y * r1 + m0.z * r2;
      const float z1 = m0.y * r0 + m0.w * r1 + m1.x * r2;
      const float z2 = m0.z * r0 + m1.x * r1 + m1.y * r2;
      R[i] = make_float4(r0, r1, r2, 0.f);
      Z[i] = make_float4(z0, z1, z2, 0.f);
      DL[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      v[0] += (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
      v[1] += (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    }
    team.template reduce<2, 0>(v, red, sel);
    rz_out = v[0];
    rr_out = v[1];
  };

  // ---- matrix-free PCG: (J^T J + D^2) delta = -J^T f -------------------
  // Iteration k reads p_old from PB[k&1], writes p = z + beta p_old to
  // PB[(k+1)&1] (own vertex) while gathering neighbours' p on the fly, so
  // the p update needs no pass or barrier of its own.
  auto pcg = [&](double rz, double rr0) -> int {
    const double tol2 = (double)opt.pcg_relative_tolerance * opt.pcg_relative_tolerance * rr0;
    double beta = 0.0;
    int k = 0;
    if (!(rr0 > 0.0)) return 0;
    for (; k < opt.pcg_max_iterations; ++k) {
      float4* __restrict__ Pold = PB[k & 1];
      float4* __restrict__ Pnew = PB[(k + 1) & 1];
      const float fb = (float)beta;
      const bool first = (k == 0);
      double v1[1] = {0.0};
      for (int i = i0; i < nv; i += stride) {
        const float4 zi = Z[i];
        float4 pi = zi;
        if (!first) { const float4 po = Pold[i]; pi.x += fb * po.x; pi.y += fb * po.y; pi.z += fb * po.z; }
        const float4 g = G[i];
        const float4 dt = DT[i];
        const float gp = g.x * pi.x + g.y * pi.y + g.z * pi.z;
        float lx = 0.f, ly = 0.f, lz = 0.f;
        const int2 a = ADJ[i];
        for (int e = 0; e < a.y; ++e) {
          const unsigned en = NB[a.x + e];
          const unsigned j = nb_index(en);
          const float m = weighted ? WT[a.x + e] : nb_mult(en);
          float4 pj = Z[j];
          if (!first) { const float4 po = Pold[j]; pj.x += fb * po.x; pj.y += fb * po.y; pj.z += fb * po.z; }
          lx += m * (pi.x - pj.x); ly += m * (pi.y - pj.y); lz += m * (pi.z - pj.z);
        }
        const float qx = g.x * gp + l2 * lx + dt.x * pi.x;
        const float qy = g.y * gp + l2 * ly + dt.y * pi.y;
        const float qz = g.z * gp + l2 * lz + dt.z * pi.z;
        Pnew[i] = make_float4(pi.x, pi.y, pi.z, 0.f);
        Q[i] = make_float4(qx, qy, qz, 0.f);
        v1[0] += (double)pi.x * qx + (double)pi.y * qy + (double)pi.z * qz;
      }
      team.template reduce<1, 0>(v1, red, sel);
      const double pq = v1[0];
      if (!(pq > 0.0)) break;
      const double alpha = rz / pq;
      const float fa = (float)alpha;
      double v2[2] = {0.0, 0.0};
      for (int i = i0; i < nv; i += stride) {
        const float4 pi = Pnew[i];
        const float4 qi = Q[i];
        float4 d = DL[i];
        d.x += fa * pi.x; d.y += fa * pi.y; d.z += fa * pi.z;
        DL[i] = d;
        float4 r = R[i];
        r.x -= fa * qi.x; r.y -= fa * qi.y; r.z -= fa * qi.z;
        R[i] = r;
        const float4 m0 = MI0[i], m1 = MI1[i];
        const float z0 = m0.x * r.x + m0.y * r.y + m0.z * r.z;
        const float z1 = m0.y * r.x + m0.w * r.y + m1.x * r.z;
        const float z2 = m0.z * r.x + m1.x * r.y + m1.y * r.z;
        Z[i] = make_float4(z0, z1, z2, 0.f);
        v2[0] += (double)r.x * z0 + (double)r.y * z1 + (double)r.z * z2;
        v2[1] += (double)r.x * r.x + (double)r.y * r.y + (double)r.z * r.z;
      }
      team.template reduce<2, 0>(v2, red, sel);
      beta = v2[0] / rz;
      rz = v2[0];
      if (v2[1] <= tol2) { ++k; break; }
    }
    return k;
  };

  // ---- model cost change, candidate cost, step and parameter norms ------
  auto eval_candidate = [&]() -> LMCandidate {
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < nv; i += stride) {
      const float4 x0 = X0[i];
      const float4 ui = U[i];
      const float4 di = DL[i];
      const float4 g = G[i];
      const float gd = g.x * di.x + g.y * di.y + g.z * di.z;
      v[0] += (double)gd * ((double)g.w + 0.5 * (double)gd);
      const float cx = ui.x + di.x, cy = ui.y + di.y, cz = ui.z + di.z;
      const float dc = lookup_f32(grid, N, x0.x + cx, x0.y + cy, x0.z + cz);
      v[1] += 0.5 * (double)dc * (double)dc;
      double mr = 0.0, ce = 0.0;
      const int2 a = ADJ[i];
      for (int e = 0; e < a.y; ++e) {
        const unsigned en = NB[a.x + e];
        const unsigned j = nb_index(en);
        const float m = weighted ? WT[a.x + e] : nb_mult(en);
        const float4 uj = U[j];
        const float4 dj = DL[j];
        const float ex = ui.x - uj.x, ey = ui.y - uj.y, ez = ui.z - uj.z;
        const float fx = di.x - dj.x, fy = di.y - dj.y, fz = di.z - dj.z;
        mr += (double)m * ((double)fx * (ex + 0.5f * fx) + (double)fy * (ey + 0.5f * fy) +
                           (double)fz * (ez + 0.5f * fz));
        const float hx = cx - (uj.x + dj.x), hy = cy - (uj.y + dj.y), hz = cz - (uj.z + dj.z);
        ce += (double)m * ((double)hx * hx + (double)hy * hy + (double)hz * hz);
      }
      v[0] += 0.5 * (double)l2 * mr;
      v[2] += 0.25 * (double)l2 * ce;
      v[3] += (double)di.x * di.x + (double)di.y * di.y + (double)di.z * di.z;
      const float px = x0.x + ui.x, py = x0.y + ui.y, pz = x0.z + ui.z;
      v[4] += (double)px * px + (double)py * py + (double)pz * pz;
    }
    team.template reduce<5, 0>(v, red, sel);
    return LMCandidate{-v[0], v[1] + v[2], sqrt(v[3]), sqrt(v[4])};
  };

  auto eval = [&](bool first, double& cv, double& ce, double& gmax) { eval_jacobian(first, cv, ce, gmax); };
  auto solve = [&](bool reuse, double radius) -> int {
    double rz, rr;
    setup_linear(reuse, radius, rz, rr);
    return pcg(rz, rr);
  };
  auto accept = [&]() {
    for (int i = i0; i < nv; i += stride) {
      float4 ui = U[i];
      const float4 di = DL[i];
      ui.x += di.x; ui.y += di.y; ui.z += di.z;
      U[i] = ui;
    }
    team.barrier();
  };
  lm_trust_region_loop(opt, eval, solve, eval_candidate, accept, team.leader() ? out : nullptr, clock_start);
}

__global__ void __launch_bounds__(1024) k_lm_solve(const LMPairDesc* __restrict__ pairs,
                                                   const int* __restrict__ plist, LMArrays A, LMOptionsDev opt,
                                                   LMResultDev* __restrict__ res, long long totalV) {
  __shared__ double red[2 * 8 * kMaxWaves];
  const int pid = plist[blockIdx.x];
  SoloTeam team;
  lm_solve_pair(pairs[pid], res + pid, A, opt, totalV, team, red);
}

// wg[b] = (team, rank, team size, -): team t solves pair plist[t].
__global__ void __launch_bounds__(1024) k_lm_solve_team(const LMPairDesc* __restrict__ pairs,
                                                        const int* __restrict__ plist, const int4* __restrict__ wg,
                                                        LMArrays A, LMOptionsDev opt, LMResultDev* __restrict__ res,
                                                        long long totalV, unsigned* cnt, double* part,
                                                        int* abort_flag, long long limit) {
  __shared__ double red[2 * 8 * kMaxWaves];
  const int4 m = wg[blockIdx.x];
  const int pid = plist[m.x];
  GridTeam team{m.y, m.z, cnt + (size_t)m.x * 32, part + (size_t)m.x * 2 * kMaxTeam * 8, abort_flag, limit, 0u, 0};
  lm_solve_pair(pairs[pid], res + pid, A, opt, totalV, team, red);
}

}  // namespace

int lm_solve_launch(const LMPairDesc* pairs_dev, const int* plist, int npairs, const LMArrays& arrays,
                    const LMOptionsDev& opt, LMResultDev* results_dev, long long total_vertices, int block_threads,
                    hipStream_t s) {
  if (npairs <= 0) return MO_OK;
  MO_CHECK(block_threads >= 64 && block_threads <= 1024 && block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 1024] (got %d)", block_threads);
  hipLaunchKernelGGL(k_lm_solve, dim3(npairs), dim3(block_threads), 0, s, pairs_dev, plist, arrays, opt,
                     results_dev, total_vertices);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

// Team size cap of the global-memory team kernels (the register/LDS team
// kernel's teams may reach kMaxTeam, the partials' stride).
int lm_team_max() { return 32; }
int lm_team_part_doubles() { return 2 * kMaxTeam * 8; }

std::mutex& team_launch_mutex() {
  static std::mutex m;
  return m;
}

int lm_team_capacity(int block_threads) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_lm_solve_team, block_threads, 0) !=
          hipSuccess)
    return 0;
  return per * cus;
}

int lm_team_launch(const LMPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg,
                   const LMArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                   long long total_vertices, int block_threads, unsigned* cnt, double* part, int* abort_flag,
                   hipStream_t s) {
  if (nwg <= 0) return MO_OK;
  MO_CHECK(block_threads >= 64 && block_threads <= 1024 && block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 1024] (got %d)", block_threads);
  int dev = 0, khz = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  long long limit = (long long)std::max(khz, 1) * 1000LL * 20;   // a 20 s bound on any one wait
  // MO_LM_TEAM_WAIT_US overrides the bound (microseconds; a negative value
  // makes the first wait give up at once: the failure path's test hook)
  if (const char* w = getenv("MO_LM_TEAM_WAIT_US")) limit = atoll(w) * (long long)std::max(khz, 1) / 1000LL;
  LMArrays a = arrays;
  LMOptionsDev o = opt;
  LMResultDev* r = results_dev;
  long long tv = total_vertices;
  void* args[] = {(void*)&pairs_dev, (void*)&plist, (void*)&wg_dev, (void*)&a, (void*)&o, (void*)&r, (void*)&tv,
                  (void*)&cnt, (void*)&part, (void*)&abort_flag, (void*)&limit};
  // MO_LM_COOP_REFUSE=1 treats the launch as refused without trying it (the
  // fallback path's test hook; the runtime refuses cooperative grids while
  // another process holds the GPU)
  if (const char* f = getenv("MO_LM_COOP_REFUSE"))
    if (f[0] == '1') return kCoopRefused;
  if (hipLaunchCooperativeKernel((const void*)k_lm_solve_team, dim3(nwg), dim3(block_threads), args, 0, s) !=
      hipSuccess) {
    (void)hipGetLastError();   // the runtime refused the cooperative grid: nothing was launched
    return kCoopRefused;
  }
  return MO_OK;
}

}  // namespace mo
