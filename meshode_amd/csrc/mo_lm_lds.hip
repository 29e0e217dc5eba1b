// Register/LDS-resident variant of the batched LM solver (same algorithm as
// mo_lm.hip: Deformer::Deform / DeformGraph, deformer.cc:18-92, 370-442,
// with Ceres' trust-region LM restated and a matrix-free block-Jacobi PCG).
//
// One workgroup of T threads solves one pair whose nv <= T * VPT vertices
// are statically owned (vertex i by thread i % T, slot i / T).  Everything
// the PCG iterates on lives in VGPRs for the whole solve: the distance
// gradient g, the direction p, q = A p, the step delta and
// d = lambda^2 deg + LM diagonal / radius (the diagonal of A apart from
// g g^T, which is also the block-Jacobi preconditioner's diagonal).  The
// one vector neighbours read (u while evaluating the Jacobian, z = M^-1 r
// inside PCG, u + delta for the candidate) is published to a
// float4-per-vertex LDS array whose w holds the vertex's adjacency word
// (LDS offset << 16 | count) for the whole solve; the PCG residual is not
// stored but recovered as r = M z from the published z.  The pair's
// adjacency is copied once into LDS as uint16 entries (neighbour << 1 |
// multiplicity - 1), each vertex's list padded to even length with a zero
// vector slot so the PCG product reads two entries per 32-bit word; pairs
// whose entries all share one multiplicity skip the per-entry weight.  The
// PCG loop touches no global memory.  Reductions are FP64, through DPP lane
// moves and one LDS exchange between waves.
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

enum GSlot { gGR = 1, gS = 2, gDG = 3, gZ = 4 };   // global slots of LMArrays::work used here

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

// x, y, z of an LDS vertex record (w, the vertex's adjacency word, stays).
typedef float mo_v3f __attribute__((ext_vector_type(3)));
__device__ inline void store_xyz(float4* p, float4 v) {
  mo_v3f t;
  t.x = v.x; t.y = v.y; t.z = v.z;
  *reinterpret_cast<mo_v3f*>(p) = t;
}

// A solve after a rejected step starts from that step (scaled) instead of
// from zero (see solve below).
constexpr bool kPcgWarmStart = true;

// An index the optimizer cannot see through: per-vertex addresses are then
// formed at each use (base + index) instead of being hoisted out of the
// solver loops, where they would hold dozens of registers for the whole
// kernel.
__device__ inline int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// The solver body, shared by the one-workgroup kernel (Team = LdsSolo: the
// workgroup owns all nv vertices, neighbours come from its LDS array) and
// the multi-workgroup kernel (Team = LdsTeam: workgroup r owns the
// contiguous chunk [lo, hi) of the pair's vertices; a neighbour in another
// chunk is read from the pair's global publication array, which every
// owner writes next to its LDS copy before the team barrier).  Entry: an
// adjacency entry type (uint16 local indices for the solo kernel, uint32
// pair indices for the team kernel).
template <int T, int VPT, class Team, class Entry>
__device__ inline void lds_solve(const LMPairDesc& P, int pid, Team& team, const LMArrays& A,
                                 const LMOptionsDev& opt, LMResultDev* __restrict__ res, long long totalV,
                                 unsigned char* smem, double* red, int* scan_buf, unsigned lds_bytes) {
  const double o_min_diag = opt.min_lm_diagonal, o_max_diag = opt.max_lm_diagonal;
  const float o_pcg_tol = opt.pcg_relative_tolerance;
  const bool o_pcg_warm = kPcgWarmStart;
  const int o_pcg_max = opt.pcg_max_iterations;
  const int tid = threadIdx.x;
  const int lo = team.lo;                  // owned vertices: lo + t + k T < hi
  const int own = team.hi - team.lo;
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
  float4* __restrict__ ZG = A.work + (size_t)gZ * totalV + vb;   // team: published vectors
  // paired layout (one workgroup per pair): a zero vector at VEC[own], and
  // every vertex's entry list padded to even length with it, so the PCG
  // product reads two entries per 32-bit LDS word
  constexpr bool kPaired = Team::kPaired;
  constexpr int kZeroSlots = kPaired ? 1 : 0;
  float4* VEC = (float4*)smem;                                      // own vertices, local index
  Entry* LCSR = (Entry*)(smem + (size_t)(own + kZeroSlots) * sizeof(float4));
  int sel = 0;
  const long long clock_start = wall_clock64();

  // ---- adjacency -> LDS (packed, compacted by a block scan) -------------
  // each own vertex's (LDS offset << 16) | count lives in VEC[il].w for the
  // whole solve (publications write x, y, z only), so it costs no register
  float uniform_m = 0.f;   // paired layout: the pair's multiplicity when all entries share it, else 0
  {
    int cnt[VPT], gst[VPT], mine = 0;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int il = opaque(tid + k * T), i = lo + il;
      cnt[k] = 0;
      gst[k] = 0;
      if (il < own) {
        const int2 a = ADJ[i];
        gst[k] = a.x;
        cnt[k] = a.y;
        mine += kPaired ? a.y + (a.y & 1) : a.y;
      }
    }
    int total;
    int off = block_exclusive_scan(mine, scan_buf, total);
    // the launch's LDS was sized from the same adjacency; a mismatch (a
    // replayed graph whose inputs' faces changed) leaves the pair unsolved
    // and is reported by the host, never written past the allocation
    if ((size_t)(own + kZeroSlots) * sizeof(float4) + (size_t)total * sizeof(Entry) > lds_bytes) return;
    int has1 = 0, has2 = 0;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      {
        const int il = opaque(tid + k * T);
        if (il < own) reinterpret_cast<float*>(VEC + il)[3] = __uint_as_float(((unsigned)off << 16) | (unsigned)cnt[k]);
      }
      for (int e = 0; e < cnt[k]; ++e) {
        const unsigned en = NB[gst[k] + e];
        has1 |= (en >> 24) == 1u;
        has2 |= (en >> 24) == 2u;
        LCSR[off + e] = (Entry)((nb_index(en) << 1) | ((en >> 24) - 1u));
      }
      if (kPaired && (cnt[k] & 1)) LCSR[off + cnt[k]] = (Entry)((unsigned)own << 1);
      off += kPaired ? cnt[k] + (cnt[k] & 1) : cnt[k];
    }
    if constexpr (kPaired) {
      const int any1 = __syncthreads_or(has1), any2 = __syncthreads_or(has2);
      uniform_m = any1 && !any2 ? 1.f : (any2 && !any1 ? 2.f : 0.f);
      if (tid == 0) VEC[own] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }

  // ---- per-vertex register state ----------------------------------------
  // g: distance gradient; p: direction; q = A p; d = lambda^2 deg + LM
  // diagonal / radius (the diagonal of A without g g^T, also the
  // preconditioner's diagonal); e: the step delta.  The PCG residual r is
  // not stored: r = M z for the z = M^-1 r held in LDS.
  float gx[VPT], gy[VPT], gz[VPT];
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
  auto gather = [&](int cs, float& sx, float& sy, float& sz) {
    sx = 0.f; sy = 0.f; sz = 0.f;
    const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
    if constexpr (kPaired) {
      // two entries per LDS word (offsets are even; an odd list ends with
      // the zero slot), two independent neighbour reads in flight
      const unsigned* __restrict__ L2 = reinterpret_cast<const unsigned*>(LCSR) + (o >> 1);
      const int n2 = (c + 1) >> 1;
      // (the entry indices pass through opaque() so each address is one
      // bit-field extract and one shift-add)
      const unsigned* const L2e = L2 + n2;
      if (uniform_m != 0.f) {
        #pragma unroll 1
        for (const unsigned* pw = L2; pw < L2e; ++pw) {
          const unsigned w = *pw;
          const float4 a = VEC[opaque((int)__builtin_amdgcn_ubfe(w, 1, 15))];
          const float4 b = VEC[opaque((int)__builtin_amdgcn_ubfe(w, 17, 15))];
          sx += a.x + b.x; sy += a.y + b.y; sz += a.z + b.z;
        }
        sx *= uniform_m; sy *= uniform_m; sz *= uniform_m;
      } else {
        #pragma unroll 1
        for (const unsigned* pw = L2; pw < L2e; ++pw) {
          const unsigned w = *pw;
          const float4 a = VEC[opaque((int)__builtin_amdgcn_ubfe(w, 1, 15))];
          const float4 b = VEC[opaque((int)__builtin_amdgcn_ubfe(w, 17, 15))];
          const float ma = (float)((w & 1u) + 1u), mb = (float)(((w >> 16) & 1u) + 1u);
          sx += ma * a.x + mb * b.x; sy += ma * a.y + mb * b.y; sz += ma * a.z + mb * b.z;
        }
      }
    } else {
      #pragma unroll 1
      for (int q = 0; q < c; ++q) {
        const unsigned w = LCSR[o + q];
        const float m = (float)((w & 1u) + 1u);
        const float4 vj = team.neighbour(VEC, ZG, w >> 1);
        sx += m * vj.x; sy += m * vj.y; sz += m * vj.z;
      }
    }
  };

  auto eval = [&](bool first, double& cv, double& ce, double& gmax) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int il = opaque(tid + k * T), i = lo + il;
      if (il < own) team.publish(VEC, ZG, il, i, U[i]);
    }
    team.sync();
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int il = opaque(tid + k * T), i = lo + il;
      if (il >= own) continue;
      const float4 x0 = X0[i];
      const float4 ui = VEC[il];
      const int cs = __float_as_int(ui.w);
      const JetF J = lookup_value_grad(grid, N, x0.x + ui.x, x0.y + ui.y, x0.z + ui.z);
      gx[k] = J.v0; gy[k] = J.v1; gz[k] = J.v2;
      const float ldk = l2 * x0.w;
      float sx = 0.f, sy = 0.f, sz = 0.f;
      double e = 0.0;
      const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
      #pragma unroll 1
      for (int q = 0; q < c; ++q) {
        const unsigned w = LCSR[o + q];
        const float m = (float)((w & 1u) + 1u);
        const float4 uj = team.neighbour(VEC, ZG, w >> 1);
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
    team.template reduce<2, 1>(v, red, sel);
    cv = v[0]; ce = v[1]; gmax = v[2];
  };

  // PCG on (J^T J + D^2 / radius) delta = -J^T f.  Each iteration publishes
  // z = M^-1 r once and forms A p as A z + beta A p_prev (p = z + beta
  // p_prev), so an iteration has two barriers: the reduction of p.Ap, and
  // the reduction of r.z and r.r, which also publishes the new z.
  //
  // After a rejected step (reuse: the same J^T J and J^T f, only the LM
  // diagonal grew) the solve starts from the rejected step e scaled by
  // alpha = (b . A e) / (e . A e), the best multiple of it in the A-norm,
  // instead of from zero; the stopping test stays |r| <= tol |b|, the same
  // absolute accuracy as a solve started from zero.
  auto solve = [&](bool reuse, double radius) -> int {
    const float finv_radius = (float)(1.0 / radius);
    const bool warm = reuse && o_pcg_warm;
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int il = opaque(tid + k * T), i = lo + il;
      px[k] = py[k] = pz[k] = 0.f;
      qx[k] = qy[k] = qz[k] = 0.f;
      if (!warm) ex[k] = ey[k] = ez[k] = 0.f;
      dx[k] = dy[k] = dz[k] = 1.f;
      if (il >= own) continue;
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
      if (warm) {
        team.publish(VEC, ZG, il, i, make_float4(ex[k], ey[k], ez[k], 0.f));
        continue;
      }
      const float4 gr = GR[i];
      const float r0 = -gr.x, r1 = -gr.y, r2 = -gr.z;
      float z0, z1, z2;
      precond(k, r0, r1, r2, z0, z1, z2);
      team.publish(VEC, ZG, il, i, make_float4(z0, z1, z2, 0.f));
      v[0] += (double)(r0 * z0 + r1 * z1 + r2 * z2);
      v[1] += (double)(r0 * r0 + r1 * r1 + r2 * r2);
    }
    if (warm) {
      // A e (into q) and the scaling alpha
      team.sync();
      double a2[2] = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int il = opaque(tid + k * T), i = lo + il;
        if (il >= own) continue;
        const float4 ei = VEC[il];
        float sx, sy, sz;
        gather(__float_as_int(ei.w), sx, sy, sz);
        const float gee = gx[k] * ei.x + gy[k] * ei.y + gz[k] * ei.z;
        qx[k] = gx[k] * gee + dx[k] * ei.x - l2 * sx;
        qy[k] = gy[k] * gee + dy[k] * ei.y - l2 * sy;
        qz[k] = gz[k] * gee + dz[k] * ei.z - l2 * sz;
        const float4 gr = GR[i];
        a2[0] -= (double)(gr.x * qx[k] + gr.y * qy[k] + gr.z * qz[k]);
        a2[1] += (double)(ex[k] * qx[k] + ey[k] * qy[k] + ez[k] * qz[k]);
      }
      team.template reduce<2, 0>(a2, red, sel);
      const float al = a2[1] > 0.0 ? (float)(a2[0] / a2[1]) : 0.f;
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int il = opaque(tid + k * T), i = lo + il;
        if (il >= own) continue;
        ex[k] *= al; ey[k] *= al; ez[k] *= al;
        const float4 gr = GR[i];
        const float r0 = -gr.x - al * qx[k], r1 = -gr.y - al * qy[k], r2 = -gr.z - al * qz[k];
        qx[k] = qy[k] = qz[k] = 0.f;
        float z0, z1, z2;
        precond(k, r0, r1, r2, z0, z1, z2);
        team.publish(VEC, ZG, il, i, make_float4(z0, z1, z2, 0.f));
        v[0] += (double)(r0 * z0 + r1 * z1 + r2 * z2);
        v[1] += (double)(r0 * r0 + r1 * r1 + r2 * r2);
        v[2] += (double)(gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
      }
    }
    team.template reduce<3, 0>(v, red, sel);   // also publishes z
    double rz = v[0];
    const double rr0 = warm ? v[2] : v[1];     // |b|^2
    if (!(rr0 > 0.0)) return 0;
    const double tol2 = (double)o_pcg_tol * o_pcg_tol * rr0;
    if (warm && v[1] <= tol2) return 0;
    float beta = 0.f;
    int it = 0;
    while (it < o_pcg_max) {
      double w1[1] = {0.0};
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int il = opaque(tid + k * T), i = lo + il;
        if (il >= own) continue;
        const float4 zi = VEC[il];
        float sx, sy, sz;
        gather(__float_as_int(zi.w), sx, sy, sz);
        const float gzz = gx[k] * zi.x + gy[k] * zi.y + gz[k] * zi.z;
        const float wx = gx[k] * gzz + dx[k] * zi.x - l2 * sx;
        const float wy = gy[k] * gzz + dy[k] * zi.y - l2 * sy;
        const float wz = gz[k] * gzz + dz[k] * zi.z - l2 * sz;
        px[k] = zi.x + beta * px[k]; py[k] = zi.y + beta * py[k]; pz[k] = zi.z + beta * pz[k];
        qx[k] = wx + beta * qx[k]; qy[k] = wy + beta * qy[k]; qz[k] = wz + beta * qz[k];
        w1[0] += (double)(px[k] * qx[k] + py[k] * qy[k] + pz[k] * qz[k]);
      }
      team.template reduce<1, 0>(w1, red, sel);
      const double pq = w1[0];
      if (!(pq > 0.0)) break;
      const float fa = (float)(rz / pq);
      double w2[2] = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        const int il = opaque(tid + k * T), i = lo + il;
        if (il >= own) continue;
        ex[k] += fa * px[k]; ey[k] += fa * py[k]; ez[k] += fa * pz[k];
        // the residual is not kept: r = M z for the z still published
        // (M = g g^T + diag(d), the preconditioner), then r -= alpha q
        const float4 zo = VEC[il];
        const float gzo = gx[k] * zo.x + gy[k] * zo.y + gz[k] * zo.z;
        const float r0 = gx[k] * gzo + dx[k] * zo.x - fa * qx[k];
        const float r1 = gy[k] * gzo + dy[k] * zo.y - fa * qy[k];
        const float r2 = gz[k] * gzo + dz[k] * zo.z - fa * qz[k];
        float z0, z1, z2;
        precond(k, r0, r1, r2, z0, z1, z2);
        team.publish(VEC, ZG, il, i, make_float4(z0, z1, z2, 0.f));
        w2[0] += (double)(r0 * z0 + r1 * z1 + r2 * z2);
        w2[1] += (double)(r0 * r0 + r1 * r1 + r2 * r2);
      }
      team.template reduce<2, 0>(w2, red, sel);   // also publishes z
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
      const int il = opaque(tid + k * T), i = lo + il;
      if (il < own) {
        const float4 ui = U[i];
        team.publish(VEC, ZG, il, i, make_float4(ui.x + ex[k], ui.y + ey[k], ui.z + ez[k], 0.f));
      }
    }
    team.sync();
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int il = opaque(tid + k * T), i = lo + il;
      if (il >= own) continue;
      const float4 x0 = X0[i];
      const float dres = GR[i].w;
      const float gd = gx[k] * ex[k] + gy[k] * ey[k] + gz[k] * ez[k];
      v[0] += (double)gd * ((double)dres + 0.5 * (double)gd);
      const float4 wv = VEC[il];
      const int cs = __float_as_int(wv.w);
      const float wx = wv.x, wy = wv.y, wz = wv.z;
      const float dc = lookup_f32(grid, N, x0.x + wx, x0.y + wy, x0.z + wz);
      v[1] += 0.5 * (double)dc * (double)dc;
      double e = 0.0;
      const int o = (int)((unsigned)cs >> 16), c = cs & 0xFFFF;
      #pragma unroll 1
      for (int q = 0; q < c; ++q) {
        const unsigned w = LCSR[o + q];
        const float m = (float)((w & 1u) + 1u);
        const float4 wj = team.neighbour(VEC, ZG, w >> 1);
        const float ax = wx - wj.x, ay = wy - wj.y, az = wz - wj.z;
        e += (double)m * ((double)ax * ax + (double)ay * ay + (double)az * az);
      }
      v[2] += 0.25 * (double)l2 * e;
      v[3] += (double)ex[k] * ex[k] + (double)ey[k] * ey[k] + (double)ez[k] * ez[k];
      const float4 ui = U[i];
      const float qx0 = x0.x + ui.x, qy0 = x0.y + ui.y, qz0 = x0.z + ui.z;
      v[4] += (double)qx0 * qx0 + (double)qy0 * qy0 + (double)qz0 * qz0;
    }
    team.template reduce<5, 0>(v, red, sel);
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
      const int il = opaque(tid + k * T), i = lo + il;
      if (il < own) {
        const float4 v = VEC[il];
        U[i] = make_float4(v.x, v.y, v.z, 0.f);
      }
    }
  };
  lm_trust_region_loop(opt, eval_tracked, solve, candidate, accept, team.leader() ? res + pid : nullptr,
                       clock_start);
}

// One workgroup per pair.
struct LdsSolo {
  static constexpr bool kPaired = true;
  int lo, hi;
  __device__ bool leader() const { return true; }
  __device__ void sync() { __syncthreads(); }
  template <int K, int M>
  __device__ void reduce(double (&v)[K + M], double* red, int& sel) { block_reduce_dpp<K, M>(v, red, sel); }
  __device__ float4 neighbour(const float4* VEC, const float4* ZG, unsigned j) const { return VEC[j]; }
  __device__ void publish(float4* VEC, float4* ZG, int il, int i, float4 v) const { store_xyz(VEC + il, v); }
};

// A team of workgroups per pair (co-resident: cooperative launch).
// Workgroup r owns [lo, hi).  The publication array ZG is written and read
// only through device-coherent stores and loads, and the team barrier and
// reductions are CoherentTeam's, so no barrier writes back or invalidates
// the L2.
struct LdsTeam {
  static constexpr bool kPaired = false;
  CoherentTeam g;
  int lo, hi;
  __device__ bool leader() const { return g.leader(); }
  __device__ void sync() { g.barrier(); }
  template <int K, int M>
  __device__ void reduce(double (&v)[K + M], double* red, int& sel) { g.template reduce_dpp<K, M>(v, red, sel); }
  __device__ float4 neighbour(const float4* VEC, const float4* ZG, unsigned j) const {
    const unsigned jl = j - (unsigned)lo;
    return jl < (unsigned)(hi - lo) ? VEC[jl] : load_coherent(ZG + j);
  }
  __device__ void publish(float4* VEC, float4* ZG, int il, int i, float4 v) const {
    store_xyz(VEC + il, v);
    store_coherent(ZG + i, v);
  }
};

// A resident grid of workgroups (as many as the chip holds at once) works
// through plist[0, count): workgroup b solves item b, then takes the next
// unsolved item from the launch's counter `next` (zero before the launch)
// until none is left.  Pairs are solved in list order (largest first), a
// new pair starts on a workgroup as soon as its previous one ends, and
// every XCD draws from the one queue; each pair's result is independent of
// which workgroup solves it.
template <int T, int VPT>
__global__ void __launch_bounds__(T) k_lm_solve_lds(const LMPairDesc* __restrict__ pairs,
                                                    const int* __restrict__ plist, int count, int* __restrict__ next,
                                                    LMArrays A, LMOptionsDev opt, LMResultDev* __restrict__ res,
                                                    long long totalV, unsigned lds_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[2 * 8 * kMaxWaves];
  __shared__ int scan_buf[kMaxWaves];
  __shared__ int s_item;
  int w = blockIdx.x;
  while (w < count) {
    const int pid = plist[w];
    const LMPairDesc P = pairs[pid];
    LdsSolo team{0, P.nv};
    lds_solve<T, VPT, LdsSolo, unsigned short>(P, pid, team, A, opt, res, totalV, smem, red, scan_buf, lds_bytes);
    __syncthreads();   // every thread is done with this pair's LDS
    if (threadIdx.x == 0) s_item = (int)gridDim.x + atomicAdd(next, 1);
    __syncthreads();
    w = s_item;
  }
}

// wg[b] = (team, rank, team size, chunk): team t solves pair plist[t];
// workgroup `rank` owns vertices [rank chunk, min(nv, (rank + 1) chunk)).
template <int T, int VPT>
__global__ void __launch_bounds__(T) k_lm_solve_lds_team(const LMPairDesc* __restrict__ pairs,
                                                         const int* __restrict__ plist, const int4* __restrict__ wg,
                                                         LMArrays A, LMOptionsDev opt, LMResultDev* __restrict__ res,
                                                         long long totalV, unsigned* cnt, double* part,
                                                         int* abort_flag, long long limit, unsigned lds_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[2 * 8 * kMaxWaves];
  __shared__ int scan_buf[kMaxWaves];
  const int4 m = wg[blockIdx.x];
  const int pid = plist[m.x];
  const LMPairDesc P = pairs[pid];
  const int lo = min(P.nv, m.y * m.w), hi = min(P.nv, lo + m.w);
  LdsTeam team{CoherentTeam{m.y, m.z, cnt + (size_t)m.x * 32, part + (size_t)m.x * 2 * kMaxTeam * 8, abort_flag, limit, 0u, 0},
               lo, hi};
  lds_solve<T, VPT, LdsTeam, unsigned>(P, pid, team, A, opt, res, totalV, smem, red, scan_buf, lds_bytes);
}

// Team launch of the register/LDS kernel: nwg workgroups (wg_dev entries)
// in one cooperative grid.
template <int T, int VPT>
int launch_team_bucket(const LMPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg, const LMArrays& A,
                       const LMOptionsDev& opt, LMResultDev* res, long long totalV, size_t lds_bytes, unsigned* cnt,
                       double* part, int* abort_flag, hipStream_t s) {
  auto kern = k_lm_solve_lds_team<T, VPT>;
  MO_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  int dev = 0, khz = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  long long limit = (long long)std::max(khz, 1) * 1000LL * 20;   // a 20 s bound on any one wait
  if (const char* w = getenv("MO_LM_TEAM_WAIT_US")) limit = atoll(w) * (long long)std::max(khz, 1) / 1000LL;
  if (const char* f = getenv("MO_LM_COOP_REFUSE"))
    if (f[0] == '1') return kCoopRefused;
  LMArrays a = A;
  LMOptionsDev o = opt;
  long long tv = totalV;
  unsigned lb = (unsigned)lds_bytes;
  void* args[] = {(void*)&pairs_dev, (void*)&plist, (void*)&wg_dev, (void*)&a,           (void*)&o,
                  (void*)&res,       (void*)&tv,    (void*)&cnt,    (void*)&part,        (void*)&abort_flag,
                  (void*)&limit,     (void*)&lb};
  if (hipLaunchCooperativeKernel((const void*)kern, dim3(nwg), dim3(T), args, (unsigned)lds_bytes, s) != hipSuccess) {
    (void)hipGetLastError();   // the runtime refused the cooperative grid: nothing was launched
    return kCoopRefused;
  }
  return MO_OK;
}

template <int T, int VPT>
int team_bucket_capacity(size_t lds_bytes) {
  auto kern = k_lm_solve_lds_team<T, VPT>;
  int dev = 0, cus = 0, per = 0;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)kern, T, lds_bytes) != hipSuccess)
    return 0;
  return per * cus;
}

// Per pair: for each team shape b, the largest chunk's adjacency entries
// (the sum of its vertices' neighbour counts; adjacency rows may have unused
// slots after merging, so row offsets do not give it), then the one-
// workgroup kernel's LDS entries (every vertex's count rounded up to even).
// out[4 p + 0..3].  One workgroup per pair.
__global__ void __launch_bounds__(256) k_chunk_entries(const LMPairDesc* __restrict__ pairs,
                                                       const int2* __restrict__ adj, int np, int c0, int c1, int c2,
                                                       int* __restrict__ out) {
  __shared__ int acc[3][kMaxTeam];
  const int p = blockIdx.x;
  const int nv = pairs[p].nv;
  const long long vb = pairs[p].vbase;
  const int caps[3] = {c0, c1, c2};
  int c[3], ch[3];
  for (int b = 0; b < 3; ++b) {
    c[b] = nv > 0 ? (nv + caps[b] - 1) / caps[b] : 0;
    ch[b] = c[b] > 0 ? (nv + c[b] - 1) / c[b] : 1;
  }
  for (int k = threadIdx.x; k < 3 * kMaxTeam; k += blockDim.x) acc[k / kMaxTeam][k % kMaxTeam] = 0;
  __syncthreads();
  __shared__ int padded;
  if (threadIdx.x == 0) padded = 0;
  int cur[3] = {-1, -1, -1}, sum[3] = {0, 0, 0}, mine = 0;
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    const int n = adj[vb + i].y;
    mine += n + (n & 1);
    for (int b = 0; b < 3; ++b) {
      if (c[b] > kMaxTeam) continue;
      const int r = i / ch[b];
      if (r != cur[b]) {
        if (cur[b] >= 0) atomicAdd(&acc[b][cur[b]], sum[b]);
        cur[b] = r;
        sum[b] = 0;
      }
      sum[b] += n;
    }
  }
  for (int b = 0; b < 3; ++b)
    if (cur[b] >= 0) atomicAdd(&acc[b][cur[b]], sum[b]);
  atomicAdd(&padded, mine);
  __syncthreads();
  if (threadIdx.x < 3) {
    const int b = threadIdx.x;
    int mx = 0;
    if (c[b] > kMaxTeam) mx = INT_MAX;   // no team of this shape holds the pair
    else
      for (int r = 0; r < c[b]; ++r) mx = max(mx, acc[b][r]);
    out[4 * p + b] = mx;
  }
  if (threadIdx.x == 3) out[4 * p + 3] = padded;
}

template <int T, int VPT>
int launch_bucket(const LMPairDesc* pairs, const int* plist, int count, int* next, const LMArrays& A,
                  const LMOptionsDev& opt, LMResultDev* res, long long totalV, size_t lds_bytes, hipStream_t s) {
  auto kern = k_lm_solve_lds<T, VPT>;
  MO_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  // one wave of workgroups: the chip's capacity for this shape and LDS size
  int dev = 0, cus = 0, per = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  MO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)kern, T, lds_bytes));
  const int nwg = std::max(1, std::min(count, std::max(per, 1) * std::max(cus, 1)));
  hipLaunchKernelGGL(kern, dim3(nwg), dim3(T), lds_bytes, s, pairs, plist, count, next, A, opt, res, totalV,
                     (unsigned)lds_bytes);
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
//      order kPref, or MO_LM_SHAPES="TxVPT,TxVPT,..." read once.
static const int kShapeT[] = {1024, 1024, 1024, 1024, 512, 512, 512, 512, 256, 256, 256, 768, 768};
static const int kShapeVPT[] = {1, 2, 4, 5, 8, 10, 12, 16, 20, 24, 32, 6, 7};
constexpr int kNumShapes = 13;
// built-in preference: 768x7 ahead of 1024x5 (5000-vertex pairs: 36.3 against
// 37.3 ms per 512-pair launch; both spill < 20 VGPRs), otherwise by capacity
static const int kPref[kNumShapes] = {0, 1, 2, 12, 3, 4, 11, 5, 6, 7, 8, 9, 10};

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
    if (p.empty()) p.assign(kPref, kPref + kNumShapes);
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

// Team shapes of the register/LDS kernel (threads x vertices per thread per
// workgroup), in preference order: the largest chunk whose LDS fits.
static const int kTeamT[] = {1024, 1024, 1024};
static const int kTeamVPT[] = {4, 3, 2};

int lm_lds_team_shapes() { return 3; }

// Preference order of the team shapes: 1024x3, then 1024x4 (only for pairs
// whose 3072-vertex team would exceed kMaxTeam), then 1024x2 (chunks whose
// adjacency does not fit the LDS otherwise) -- 1024x3 solved 12k-100k vertex
// pairs as fast as 1024x4 and 4% faster on 41k-vertex icospheres (its
// register allocation spills 6 values against 33) -- or
// MO_LM_TEAM_VPT="v,v,..." (vertices per thread, e.g. "2,4") read once.
const std::vector<int>& lm_lds_team_order() {
  static const std::vector<int> order = [] {
    std::vector<int> o;
    if (const char* env = getenv("MO_LM_TEAM_VPT")) {
      std::string e(env);
      size_t pos = 0;
      while (pos <= e.size()) {
        const size_t comma = e.find(',', pos);
        const int v = atoi(e.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos).c_str());
        for (int b = 0; b < 3; ++b)
          if (kTeamVPT[b] == v) o.push_back(b);
        if (comma == std::string::npos) break;
        pos = comma + 1;
      }
    }
    if (o.empty()) o = {1, 0, 2};
    return o;
  }();
  return order;
}

int lm_lds_team_max() { return kMaxTeam; }

int lm_lds_team_shape(int b, int* threads, int* vpt) {
  MO_CHECK(b >= 0 && b < 3, MO_ERR_ARG, "team shape %d out of range", b);
  *threads = kTeamT[b];
  *vpt = kTeamVPT[b];
  return MO_OK;
}

size_t lm_lds_team_bytes(int chunk, long long entries) {
  return (((size_t)chunk * 16 + (size_t)entries * 4) + 15) & ~(size_t)15;
}

int lm_lds_team_capacity(int b, size_t lds_bytes) {
  switch (b) {
    case 0: return team_bucket_capacity<1024, 4>(lds_bytes);
    case 1: return team_bucket_capacity<1024, 3>(lds_bytes);
    case 2: return team_bucket_capacity<1024, 2>(lds_bytes);
    default: return 0;
  }
}

int lm_lds_chunk_entries(const LMPairDesc* pairs_dev, const int2* adj, int np, int* out_dev, hipStream_t s) {
  if (np <= 0) return MO_OK;
  hipLaunchKernelGGL(k_chunk_entries, dim3(np), dim3(256), 0, s, pairs_dev, adj, np,
                     kTeamT[0] * kTeamVPT[0], kTeamT[1] * kTeamVPT[1], kTeamT[2] * kTeamVPT[2], out_dev);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

int lm_lds_team_launch(int b, const LMPairDesc* pairs, const int* plist, const int4* wg_dev, int nwg,
                       const LMArrays& A, const LMOptionsDev& opt, LMResultDev* res, long long totalV,
                       size_t lds_bytes, unsigned* cnt, double* part, int* abort_flag, hipStream_t s) {
  if (nwg <= 0) return MO_OK;
  switch (b) {
    case 0: return launch_team_bucket<1024, 4>(pairs, plist, wg_dev, nwg, A, opt, res, totalV, lds_bytes, cnt, part,
                                               abort_flag, s);
    case 1: return launch_team_bucket<1024, 3>(pairs, plist, wg_dev, nwg, A, opt, res, totalV, lds_bytes, cnt, part,
                                               abort_flag, s);
    case 2: return launch_team_bucket<1024, 2>(pairs, plist, wg_dev, nwg, A, opt, res, totalV, lds_bytes, cnt, part,
                                               abort_flag, s);
    default: return fail(MO_ERR_ARG, "bad register/LDS team shape %d", b);
  }
}

size_t lm_lds_bytes(int nv, long long padded_entries) {
  return (((size_t)nv * 16 + 16 + (size_t)padded_entries * 2) + 15) & ~(size_t)15;
}

int lm_lds_launch(int bucket, const LMPairDesc* pairs, const int* plist, int count, int* next, const LMArrays& A,
                  const LMOptionsDev& opt, LMResultDev* res, long long totalV, size_t lds_bytes, hipStream_t s) {
  if (count <= 0) return MO_OK;
  switch (bucket) {
#define MO_SHAPE(ID, T, V) \
    case ID: return launch_bucket<T, V>(pairs, plist, count, next, A, opt, res, totalV, lds_bytes, s);
    MO_SHAPE(0, 1024, 1) MO_SHAPE(1, 1024, 2) MO_SHAPE(2, 1024, 4) MO_SHAPE(3, 1024, 5)
    MO_SHAPE(4, 512, 8) MO_SHAPE(5, 512, 10) MO_SHAPE(6, 512, 12) MO_SHAPE(7, 512, 16)
    MO_SHAPE(8, 256, 20) MO_SHAPE(9, 256, 24) MO_SHAPE(10, 256, 32) MO_SHAPE(11, 768, 6) MO_SHAPE(12, 768, 7)
#undef MO_SHAPE
    default: return fail(MO_ERR_ARG, "bad LDS solver shape %d", bucket);
  }
}

}  // namespace mo
