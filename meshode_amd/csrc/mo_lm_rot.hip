// Levenberg-Marquardt solve of Deformer::DeformWithRot (deformer.cc:94-171)
// on gfx950: one workgroup per problem, the whole trust-region loop in one
// persistent kernel (the loop of mo_lm_util.h).
//
// Unknowns per vertex i: the position p_i = x0_i + u_i and an angle-axis
// rotation w_i (zero at the start, deformer.cc:114).  Residuals:
//   DistanceLoss   d(p_i)                                  (distanceloss.h:6-25)
//   EdgeLossWithRot, per face half-edge h = (a, b) = (F[f][j], F[f][j+1]):
//     lambda (rotate(w_a, p_a - p_b) - v_h),  v_h = x0_a - x0_b,
//     w_a - w_b                                          (edgeloss.h:64-90)
// with the rotation and its derivatives of mo_rotate.h.  The Jacobian of
// half-edge h has the blocks lambda A_a (p_a), -lambda A_a (p_b),
// lambda G_h (w_a) for the first three residuals (A_a = dq/dp at w_a,
// G_h = dq/dw at (w_a, p_a - p_b)) and I (w_a), -I (w_b) for the last three.
//
// Per LM step, (J^T J + D^2) delta = -J^T f is solved by PCG with FP32
// vectors and FP64 reductions; J^T J x is formed matrix-free in two passes
// (per half-edge: t = lambda (A_a (x_pa - x_pb) + G_h x_wa) and
// s = x_wa - x_wb and their transposed images; per vertex: the gather over
// its outgoing and incoming half-edges), preconditioned by the inverse of
// each vertex's 6x6 diagonal block of J^T J + D^2.  Jacobi scaling, LM
// diagonal and trust-region decisions are those of the shared loop.
#include <algorithm>
#include <cstdlib>
#include "mo_common.h"
#include "mo_lm.h"
#include "mo_lm_rot.h"
#include "mo_lm_util.h"
#include "mo_lookup.h"
#include "mo_rotate.h"

namespace mo {

namespace {

enum VSlot {
  vG = 0, vA0, vA1, vA2, vSp, vSw, vGRp, vGRw, vDGp, vDGw, vDTp, vDTw,
  vM0, vM1, vM2, vM3, vM4, vM5, vRp, vRw, vZp, vZw, vPp, vPw, vQp, vQw, vDLp, vDLw, vNumSlots
};
enum HSlot { hG0 = 0, hG1, hG2, hSig, hC, hD, hE, hNumSlots };
static_assert(vNumSlots == kRotVSlots && hNumSlots == kRotHSlots, "slot counts are part of the host contract");

__device__ inline float3 f3(const float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ inline float4 f4(const float3 v) { return make_float4(v.x, v.y, v.z, 0.f); }
__device__ inline float3 add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ inline double dot3(float3 a, float3 b) { return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z; }
// M v and M^T v for a matrix stored as three float4 rows
__device__ inline float3 mv(const float4 r0, const float4 r1, const float4 r2, float3 v) {
  return make_float3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
                     r2.x * v.x + r2.y * v.y + r2.z * v.z);
}
__device__ inline float3 mtv(const float4 r0, const float4 r1, const float4 r2, float3 v) {
  return make_float3(r0.x * v.x + r1.x * v.y + r2.x * v.z, r0.y * v.x + r1.y * v.y + r2.y * v.z,
                     r0.z * v.x + r1.z * v.y + r2.z * v.z);
}
__device__ inline float3 absmax3(float3 a) {
  return make_float3(fabsf(a.x), fabsf(a.y), fabsf(a.z));
}

// packed upper triangle of a symmetric 6x6 matrix
__device__ inline int sym6(int r, int c) {
  if (r > c) { const int t = r; r = c; c = t; }
  return r * 6 - (r * (r - 1)) / 2 + (c - r);
}

// Inverse of an SPD 6x6 matrix (FP64 Cholesky); false when not SPD.
__device__ inline bool spd6_inverse(double (&M)[6][6], float (&out)[21]) {
  double L[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = M[i][j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  // inverse column by column: L L^T x = e_c
  for (int c = 0; c < 6; ++c) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
      y[i] = s / L[i][i];
    }
    double x[6];
    for (int i = 5; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
      x[i] = s / L[i][i];
    }
    for (int r = 0; r <= c; ++r) out[sym6(r, c)] = (float)x[r];
  }
  return true;
}

template <class Team>
__device__ inline void lm_rot_pair(const RotPairDesc P, LMResultDev* out, const RotArrays& A,
                                   const LMOptionsDev& opt, long long totalV, long long totalH, Team& team,
                                   double* red) {
  const int T = blockDim.x, tid = threadIdx.x;
  const int i0 = team.rank() * T + tid, stride = team.size() * T;
  const int nv = P.nv, nh = P.nh;
  const float lam = P.lambda, l2 = lam * lam;
  const float* __restrict__ grid = P.grid;
  const int N = P.N;
  const long long vb = P.vbase, hb = P.hbase;
  const float4* __restrict__ X0 = A.x0 + vb;
  float4* __restrict__ U = A.u + vb;
  float4* __restrict__ WR = A.w + vb;
  const int4* __restrict__ ADJ = A.vadj + vb;
  const int* __restrict__ VL = A.vlist + P.lbase;
  const int2* __restrict__ HAB = A.hab + hb;
  float4* __restrict__ V[vNumSlots];
  for (int k = 0; k < vNumSlots; ++k) V[k] = A.vwork + (size_t)k * totalV + vb;
  float4* __restrict__ H[hNumSlots];
  for (int k = 0; k < hNumSlots; ++k) H[k] = A.hwork + (size_t)k * totalH + hb;
  int sel = 0;
  const long long clock_start = wall_clock64();

  // J^T J diagonal block of vertex i (6x6, p then w), without D^2
  auto vertex_block = [&](int i, double (&M)[6][6]) {
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) M[r][c] = 0.0;
    const float4 g = V[vG][i];
    const float gv[3] = {g.x, g.y, g.z};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M[r][c] += (double)gv[r] * gv[c];
    const int4 adj = ADJ[i];
    const float4 a0 = V[vA0][i], a1 = V[vA1][i], a2 = V[vA2][i];
    const float Ai[9] = {a0.x, a0.y, a0.z, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z};
    // outgoing: l2 A_i^T A_i (p,p), l2 A_i^T G (p,w), l2 G^T G + I (w,w)
    for (int k = 0; k < adj.y; ++k) {
      const int h = VL[adj.x + k];
      const float4 q0 = H[hG0][h], q1 = H[hG1][h], q2 = H[hG2][h];
      const float Gm[9] = {q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double pp = 0, pw = 0, ww = 0;
          for (int m = 0; m < 3; ++m) {
            pp += (double)Ai[3 * m + r] * Ai[3 * m + c];
            pw += (double)Ai[3 * m + r] * Gm[3 * m + c];
            ww += (double)Gm[3 * m + r] * Gm[3 * m + c];
          }
          M[r][c] += l2 * pp;
          M[r][3 + c] += l2 * pw;
          M[3 + c][r] += l2 * pw;
          M[3 + r][3 + c] += l2 * ww + (r == c ? 1.0 : 0.0);
        }
    }
    // incoming: l2 A_a^T A_a (p,p), I (w,w)
    for (int k = 0; k < adj.z; ++k) {
      const int h = VL[adj.x + adj.y + k];
      const int a = HAB[h].x;
      const float4 b0 = V[vA0][a], b1 = V[vA1][a], b2 = V[vA2][a];
      const float Aa[9] = {b0.x, b0.y, b0.z, b1.x, b1.y, b1.z, b2.x, b2.y, b2.z};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double pp = 0;
          for (int m = 0; m < 3; ++m) pp += (double)Aa[3 * m + r] * Aa[3 * m + c];
          M[r][c] += l2 * pp;
        }
      for (int r = 0; r < 3; ++r) M[3 + r][3 + r] += 1.0;
    }
  };

  auto apply_minv = [&](int i, float3 rp, float3 rw, float3& zp, float3& zw) {
    float m[24];
    for (int k = 0; k < 6; ++k) {
      const float4 t = V[vM0 + k][i];
      m[4 * k] = t.x; m[4 * k + 1] = t.y; m[4 * k + 2] = t.z; m[4 * k + 3] = t.w;
    }
    const float r[6] = {rp.x, rp.y, rp.z, rw.x, rw.y, rw.z};
    float z[6];
    for (int a = 0; a < 6; ++a) {
      float s = 0.f;
      for (int b = 0; b < 6; ++b) s += m[sym6(a, b)] * r[b];
      z[a] = s;
    }
    zp = make_float3(z[0], z[1], z[2]);
    zw = make_float3(z[3], z[4], z[5]);
  };

  // ---- residuals, Jacobian blocks, gradient J^T f, cost -----------------
  auto eval_jacobian = [&](bool first, double& cv, double& ce, double& gmax) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < nv; i += stride) {
      const float4 x0 = X0[i], ui = U[i], wi = WR[i];
      const JetF J = lookup_jet(grid, N, x0.x + ui.x, x0.y + ui.y, x0.z + ui.z);
      V[vG][i] = make_float4(J.v0, J.v1, J.v2, J.a);
      const float w[3] = {wi.x, wi.y, wi.z};
      float Am[9];
      rotation_point_jacobian(w, Am);
      V[vA0][i] = make_float4(Am[0], Am[1], Am[2], 0.f);
      V[vA1][i] = make_float4(Am[3], Am[4], Am[5], 0.f);
      V[vA2][i] = make_float4(Am[6], Am[7], Am[8], 0.f);
      v[0] += 0.5 * (double)J.a * (double)J.a;
    }
    team.barrier();
    for (int h = i0; h < nh; h += stride) {
      const int2 ab = HAB[h];
      const float4 xa = X0[ab.x], xb = X0[ab.y], ua = U[ab.x], ub = U[ab.y], wa = WR[ab.x], wb = WR[ab.y];
      const float3 vh = make_float3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
      const float e[3] = {vh.x + (ua.x - ub.x), vh.y + (ua.y - ub.y), vh.z + (ua.z - ub.z)};
      const float w[3] = {wa.x, wa.y, wa.z};
      float q[3], G[9];
      rotate_point_dw(w, e, q, G);
      const float3 r = make_float3(lam * (q[0] - vh.x), lam * (q[1] - vh.y), lam * (q[2] - vh.z));
      const float3 sg = make_float3(wa.x - wb.x, wa.y - wb.y, wa.z - wb.z);
      const float4 g0 = make_float4(G[0], G[1], G[2], r.x), g1 = make_float4(G[3], G[4], G[5], r.y),
                   g2 = make_float4(G[6], G[7], G[8], r.z);
      H[hG0][h] = g0;
      H[hG1][h] = g1;
      H[hG2][h] = g2;
      H[hSig][h] = f4(sg);
      const float3 c = mul(mtv(V[vA0][ab.x], V[vA1][ab.x], V[vA2][ab.x], r), lam);
      const float3 d = add(mul(mtv(g0, g1, g2, r), lam), sg);
      H[hC][h] = f4(c);
      H[hD][h] = f4(d);
      v[1] += 0.5 * (dot3(r, r) + dot3(sg, sg));
    }
    team.barrier();
    for (int i = i0; i < nv; i += stride) {
      const float4 g = V[vG][i];
      const int4 adj = ADJ[i];
      float3 gp = make_float3(g.w * g.x, g.w * g.y, g.w * g.z), gw = make_float3(0.f, 0.f, 0.f);
      for (int k = 0; k < adj.y; ++k) {
        const int h = VL[adj.x + k];
        gp = add(gp, f3(H[hC][h]));
        gw = add(gw, f3(H[hD][h]));
      }
      for (int k = 0; k < adj.z; ++k) {
        const int h = VL[adj.x + adj.y + k];
        gp = sub(gp, f3(H[hC][h]));
        gw = sub(gw, f3(H[hSig][h]));
      }
      V[vGRp][i] = f4(gp);
      V[vGRw][i] = f4(gw);
      const float3 m1 = absmax3(gp), m2 = absmax3(gw);
      const float gm = fmaxf(fmaxf(fmaxf(m1.x, m1.y), fmaxf(m1.z, m2.x)), fmaxf(m2.y, m2.z));
      v[2] = (double)gm > v[2] ? (double)gm : v[2];
      if (first) {
        double M[6][6];
        vertex_block(i, M);
        float s[6];
        for (int k = 0; k < 6; ++k) s[k] = (float)(1.0 / (1.0 + sqrt(M[k][k])));
        V[vSp][i] = make_float4(s[0], s[1], s[2], 0.f);
        V[vSw][i] = make_float4(s[3], s[4], s[5], 0.f);
      }
    }
    team.template reduce<2, 1>(v, red, sel);
    cv = v[0]; ce = v[1]; gmax = v[2];
  };

  // ---- LM diagonal, 6x6 block preconditioner, PCG start ----------------
  auto setup_linear = [&](bool reuse, double radius, double& rz_out, double& rr_out) {
    double v[2] = {0.0, 0.0};
    const double inv_radius = 1.0 / radius;
    for (int i = i0; i < nv; i += stride) {
      double M[6][6];
      vertex_block(i, M);
      const float4 sp = V[vSp][i], sw = V[vSw][i];
      const float s[6] = {sp.x, sp.y, sp.z, sw.x, sw.y, sw.z};
      float dg[6];
      if (!reuse) {
        for (int k = 0; k < 6; ++k) {
          double t = (double)s[k] * s[k] * M[k][k];
          t = t < opt.min_lm_diagonal ? opt.min_lm_diagonal : (t > opt.max_lm_diagonal ? opt.max_lm_diagonal : t);
          dg[k] = (float)t;
        }
        V[vDGp][i] = make_float4(dg[0], dg[1], dg[2], 0.f);
        V[vDGw][i] = make_float4(dg[3], dg[4], dg[5], 0.f);
      } else {
        const float4 a = V[vDGp][i], b = V[vDGw][i];
        dg[0] = a.x; dg[1] = a.y; dg[2] = a.z; dg[3] = b.x; dg[4] = b.y; dg[5] = b.z;
      }
      float dt[6];
      for (int k = 0; k < 6; ++k) {
        const double t = (double)dg[k] / ((double)s[k] * s[k]) * inv_radius;
        dt[k] = (float)(t < 1e30 ? t : 1e30);
        M[k][k] += dt[k];
      }
      V[vDTp][i] = make_float4(dt[0], dt[1], dt[2], 0.f);
      V[vDTw][i] = make_float4(dt[3], dt[4], dt[5], 0.f);
      float inv[21];
      if (!spd6_inverse(M, inv)) {
        for (int k = 0; k < 21; ++k) inv[k] = 0.f;
        for (int k = 0; k < 6; ++k) inv[sym6(k, k)] = M[k][k] > 0.0 ? (float)(1.0 / M[k][k]) : 0.f;
      }
      for (int k = 0; k < 5; ++k)
        V[vM0 + k][i] = make_float4(inv[4 * k], inv[4 * k + 1], inv[4 * k + 2], inv[4 * k + 3]);
      V[vM5][i] = make_float4(inv[20], 0.f, 0.f, 0.f);
      const float3 rp = mul(f3(V[vGRp][i]), -1.f), rw = mul(f3(V[vGRw][i]), -1.f);
      float3 zp, zw;
      apply_minv(i, rp, rw, zp, zw);
      V[vRp][i] = f4(rp);
      V[vRw][i] = f4(rw);
      V[vZp][i] = f4(zp);
      V[vZw][i] = f4(zw);
      V[vPp][i] = f4(zp);
      V[vPw][i] = f4(zw);
      V[vDLp][i] = make_float4(0.f, 0.f, 0.f, 0.f);
      V[vDLw][i] = make_float4(0.f, 0.f, 0.f, 0.f);
      v[0] += dot3(rp, zp) + dot3(rw, zw);
      v[1] += dot3(rp, rp) + dot3(rw, rw);
    }
    team.template reduce<2, 0>(v, red, sel);
    rz_out = v[0];
    rr_out = v[1];
  };

  // ---- matrix-free PCG ---------------------------------------------------
  auto pcg = [&](double rz, double rr0) -> int {
    const double tol2 = (double)opt.pcg_relative_tolerance * opt.pcg_relative_tolerance * rr0;
    if (!(rr0 > 0.0)) return 0;
    int k = 0;
    for (; k < opt.pcg_max_iterations; ++k) {
      for (int h = i0; h < nh; h += stride) {
        const int2 ab = HAB[h];
        const float3 pa = f3(V[vPp][ab.x]), pb = f3(V[vPp][ab.y]);
        const float3 wa = f3(V[vPw][ab.x]), wb = f3(V[vPw][ab.y]);
        const float4 a0 = V[vA0][ab.x], a1 = V[vA1][ab.x], a2 = V[vA2][ab.x];
        const float4 g0 = H[hG0][h], g1 = H[hG1][h], g2 = H[hG2][h];
        const float3 t = mul(add(mv(a0, a1, a2, sub(pa, pb)), mv(g0, g1, g2, wa)), lam);
        const float3 s = sub(wa, wb);
        H[hC][h] = f4(mul(mtv(a0, a1, a2, t), lam));
        H[hD][h] = f4(add(mul(mtv(g0, g1, g2, t), lam), s));
        H[hE][h] = f4(s);
      }
      team.barrier();
      double v1[1] = {0.0};
      for (int i = i0; i < nv; i += stride) {
        const float3 pp = f3(V[vPp][i]), pw = f3(V[vPw][i]);
        const float4 g = V[vG][i];
        const float gd = g.x * pp.x + g.y * pp.y + g.z * pp.z;
        const float4 dtp = V[vDTp][i], dtw = V[vDTw][i];
        float3 qp = make_float3(g.x * gd + dtp.x * pp.x, g.y * gd + dtp.y * pp.y, g.z * gd + dtp.z * pp.z);
        float3 qw = make_float3(dtw.x * pw.x, dtw.y * pw.y, dtw.z * pw.z);
        const int4 adj = ADJ[i];
        for (int e = 0; e < adj.y; ++e) {
          const int h = VL[adj.x + e];
          qp = add(qp, f3(H[hC][h]));
          qw = add(qw, f3(H[hD][h]));
        }
        for (int e = 0; e < adj.z; ++e) {
          const int h = VL[adj.x + adj.y + e];
          qp = sub(qp, f3(H[hC][h]));
          qw = sub(qw, f3(H[hE][h]));
        }
        V[vQp][i] = f4(qp);
        V[vQw][i] = f4(qw);
        v1[0] += dot3(pp, qp) + dot3(pw, qw);
      }
      team.template reduce<1, 0>(v1, red, sel);
      const double pq = v1[0];
      if (!(pq > 0.0)) break;
      const float fa = (float)(rz / pq);
      double v2[2] = {0.0, 0.0};
      for (int i = i0; i < nv; i += stride) {
        const float3 pp = f3(V[vPp][i]), pw = f3(V[vPw][i]);
        V[vDLp][i] = f4(add(f3(V[vDLp][i]), mul(pp, fa)));
        V[vDLw][i] = f4(add(f3(V[vDLw][i]), mul(pw, fa)));
        const float3 rp = sub(f3(V[vRp][i]), mul(f3(V[vQp][i]), fa));
        const float3 rw = sub(f3(V[vRw][i]), mul(f3(V[vQw][i]), fa));
        V[vRp][i] = f4(rp);
        V[vRw][i] = f4(rw);
        float3 zp, zw;
        apply_minv(i, rp, rw, zp, zw);
        V[vZp][i] = f4(zp);
        V[vZw][i] = f4(zw);
        v2[0] += dot3(rp, zp) + dot3(rw, zw);
        v2[1] += dot3(rp, rp) + dot3(rw, rw);
      }
      team.template reduce<2, 0>(v2, red, sel);
      const float fb = (float)(v2[0] / rz);
      rz = v2[0];
      if (v2[1] <= tol2) { ++k; break; }
      for (int i = i0; i < nv; i += stride) {
        V[vPp][i] = f4(add(f3(V[vZp][i]), mul(f3(V[vPp][i]), fb)));
        V[vPw][i] = f4(add(f3(V[vZw][i]), mul(f3(V[vPw][i]), fb)));
      }
      team.barrier();
    }
    return k;
  };

  // ---- model cost change, candidate cost, step and parameter norms ------
  auto eval_candidate = [&]() -> LMCandidate {
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < nv; i += stride) {
      const float4 x0 = X0[i], ui = U[i], wi = WR[i];
      const float3 dp = f3(V[vDLp][i]), dw = f3(V[vDLw][i]);
      const float4 g = V[vG][i];
      const float gd = g.x * dp.x + g.y * dp.y + g.z * dp.z;
      v[0] += (double)gd * ((double)g.w + 0.5 * (double)gd);
      const float dc = lookup_f32(grid, N, x0.x + (ui.x + dp.x), x0.y + (ui.y + dp.y), x0.z + (ui.z + dp.z));
      v[1] += 0.5 * (double)dc * (double)dc;
      v[3] += dot3(dp, dp) + dot3(dw, dw);
      const float3 x = make_float3(x0.x + ui.x, x0.y + ui.y, x0.z + ui.z);
      v[4] += dot3(x, x) + dot3(f3(wi), f3(wi));
    }
    for (int h = i0; h < nh; h += stride) {
      const int2 ab = HAB[h];
      const float3 dpa = f3(V[vDLp][ab.x]), dpb = f3(V[vDLp][ab.y]);
      const float3 dwa = f3(V[vDLw][ab.x]), dwb = f3(V[vDLw][ab.y]);
      const float4 a0 = V[vA0][ab.x], a1 = V[vA1][ab.x], a2 = V[vA2][ab.x];
      const float4 g0 = H[hG0][h], g1 = H[hG1][h], g2 = H[hG2][h];
      const float3 r = make_float3(g0.w, g1.w, g2.w), sg = f3(H[hSig][h]);
      const float3 m = mul(add(mv(a0, a1, a2, sub(dpa, dpb)), mv(g0, g1, g2, dwa)), lam);
      const float3 ms = sub(dwa, dwb);
      v[0] += dot3(m, make_float3(r.x + 0.5f * m.x, r.y + 0.5f * m.y, r.z + 0.5f * m.z)) +
              dot3(ms, make_float3(sg.x + 0.5f * ms.x, sg.y + 0.5f * ms.y, sg.z + 0.5f * ms.z));
      const float4 xa = X0[ab.x], xb = X0[ab.y], ua = U[ab.x], ub = U[ab.y], wa = WR[ab.x], wb = WR[ab.y];
      const float3 vh = make_float3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
      const float e[3] = {vh.x + ((ua.x + dpa.x) - (ub.x + dpb.x)), vh.y + ((ua.y + dpa.y) - (ub.y + dpb.y)),
                          vh.z + ((ua.z + dpa.z) - (ub.z + dpb.z))};
      const float w[3] = {wa.x + dwa.x, wa.y + dwa.y, wa.z + dwa.z};
      float q[3];
      rotate_point(w, e, q);
      const float3 rc = make_float3(lam * (q[0] - vh.x), lam * (q[1] - vh.y), lam * (q[2] - vh.z));
      const float3 sc = make_float3(w[0] - (wb.x + dwb.x), w[1] - (wb.y + dwb.y), w[2] - (wb.z + dwb.z));
      v[2] += 0.5 * (dot3(rc, rc) + dot3(sc, sc));
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
      U[i] = f4(add(f3(U[i]), f3(V[vDLp][i])));
      WR[i] = f4(add(f3(WR[i]), f3(V[vDLw][i])));
    }
    team.barrier();
  };
  lm_trust_region_loop(opt, eval, solve, eval_candidate, accept, team.leader() ? out : nullptr, clock_start);
}

__global__ void __launch_bounds__(256) k_lm_rot(const RotPairDesc* __restrict__ pairs, RotArrays A,
                                                 LMOptionsDev opt, LMResultDev* __restrict__ res, long long totalV,
                                                 long long totalH) {
  __shared__ double red[2 * 8 * kMaxWaves];
  const int pid = blockIdx.x;
  SoloTeam team;
  lm_rot_pair(pairs[pid], res + pid, A, opt, totalV, totalH, team, red);
}

// wg[b] = (team, rank, team size, -): team t solves problem plist[t].
__global__ void __launch_bounds__(256) k_lm_rot_team(const RotPairDesc* __restrict__ pairs,
                                                      const int* __restrict__ plist, const int4* __restrict__ wg,
                                                      RotArrays A, LMOptionsDev opt, LMResultDev* __restrict__ res,
                                                      long long totalV, long long totalH, unsigned* cnt, double* part,
                                                      int* abort_flag, long long limit) {
  __shared__ double red[2 * 8 * kMaxWaves];
  const int4 m = wg[blockIdx.x];
  const int pid = plist[m.x];
  GridTeam team{m.y, m.z, cnt + (size_t)m.x * 32, part + (size_t)m.x * 2 * kMaxTeam * 8, abort_flag, limit, 0u, 0};
  lm_rot_pair(pairs[pid], res + pid, A, opt, totalV, totalH, team, red);
}

}  // namespace

int lm_rot_launch(const RotPairDesc* pairs_dev, int npairs, const RotArrays& arrays, const LMOptionsDev& opt,
                  LMResultDev* results_dev, long long total_vertices, long long total_half_edges, int block_threads,
                  hipStream_t s) {
  if (npairs <= 0) return MO_OK;
  MO_CHECK(block_threads >= 64 && block_threads <= 256 && block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 256] (got %d)", block_threads);
  hipLaunchKernelGGL(k_lm_rot, dim3(npairs), dim3(block_threads), 0, s, pairs_dev, arrays, opt, results_dev,
                     total_vertices, total_half_edges);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

int lm_rot_team_capacity(int block_threads) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_lm_rot_team, block_threads, 0) != hipSuccess)
    return 0;
  return per * cus;
}

int lm_rot_team_launch(const RotPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg,
                       const RotArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                       long long total_vertices, long long total_half_edges, int block_threads, unsigned* cnt,
                       double* part, int* abort_flag, hipStream_t s) {
  if (nwg <= 0) return MO_OK;
  MO_CHECK(block_threads >= 64 && block_threads <= 256 && block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 256] (got %d)", block_threads);
  int dev = 0, khz = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  long long limit = (long long)std::max(khz, 1) * 1000LL * 20;   // a 20 s bound on any one wait
  if (const char* w = getenv("MO_LM_TEAM_WAIT_US")) limit = atoll(w) * (long long)std::max(khz, 1) / 1000LL;
  RotArrays a = arrays;
  LMOptionsDev o = opt;
  LMResultDev* r = results_dev;
  long long tv = total_vertices, th = total_half_edges;
  void* args[] = {(void*)&pairs_dev, (void*)&plist, (void*)&wg_dev, (void*)&a, (void*)&o, (void*)&r, (void*)&tv,
                  (void*)&th, (void*)&cnt, (void*)&part, (void*)&abort_flag, (void*)&limit};
  // MO_LM_COOP_REFUSE=1 treats the launch as refused without trying it (the
  // fallback path's test hook; the runtime refuses cooperative grids while
  // another process holds the GPU)
  if (const char* f = getenv("MO_LM_COOP_REFUSE"))
    if (f[0] == '1') return kCoopRefused;
  if (hipLaunchCooperativeKernel((const void*)k_lm_rot_team, dim3(nwg), dim3(block_threads), args, 0, s) !=
      hipSuccess) {
    (void)hipGetLastError();   // the runtime refused the cooperative grid: nothing was launched
    return kCoopRefused;
  }
  return MO_OK;
}

}  // namespace mo
