// Device trilinear distance lookup with the exact arithmetic of
// UniformGrid::DistanceFloat<T> (uniformgrid.cc:85-150), T = float and
// T = Jet<float,3> (Ceres' dual number: (f*g).v = f.a*g.v + f.v*g.a).
//
// Translation units including this header are compiled with
// -ffp-contract=off so that every operation is one FP32 rounding in the
// order written, which makes value and derivative bit-identical to the
// reference's x86 build.
#pragma once

#include "mo_common.h"

namespace mo {

struct JetF {
  float a, v0, v1, v2;
};

__device__ inline JetF jconst(float s) { return JetF{s, 0.0f, 0.0f, 0.0f}; }
__device__ inline JetF jmul(const JetF& f, const JetF& g) {
  float x0 = f.a * g.v0, y0 = f.v0 * g.a;
  float x1 = f.a * g.v1, y1 = f.v1 * g.a;
  float x2 = f.a * g.v2, y2 = f.v2 * g.a;
  return JetF{f.a * g.a, x0 + y0, x1 + y1, x2 + y2};
}
__device__ inline JetF jadd(const JetF& f, const JetF& g) {
  return JetF{f.a + g.a, f.v0 + g.v0, f.v1 + g.v1, f.v2 + g.v2};
}
__device__ inline JetF jsub(const JetF& f, const JetF& g) {
  return JetF{f.a - g.a, f.v0 - g.v0, f.v1 - g.v1, f.v2 - g.v2};
}
__device__ inline JetF jneg(const JetF& f) { return JetF{-f.a, -f.v0, -f.v1, -f.v2}; }

// Scalar DistanceFloat<float>.  grid: FP32 voxels (the reference's FP64
// voxels rounded to float, which is what (T)voxel does at read time).
__device__ inline float lookup_f32(const float* __restrict__ grid, int N, float p0, float p1, float p2) {
  const float fN = (float)N;
  const int px = trunc_i32f(p0 * fN), py = trunc_i32f(p1 * fN), pz = trunc_i32f(p2 * fN);
  if (px < 0 || py < 0 || pz < 0 || px >= N - 1 || py >= N - 1 || pz >= N - 1) {
    const float edge = (float)(N - 1 - 1e-3);
    float l = 0.0f;
    if (px < 0) l = l + (-p0) * fN;
    else if (px >= N) l = l + (p0 * fN - edge);
    if (py < 0) l = l + (-p1) * fN;
    else if (py >= N) l = l + (p1 * fN - edge);
    if (pz < 0) l = l + (-p2) * fN;
    else if (pz >= N) l = l + (p2 * fN - edge);
    return l;
  }
  const float wx = p0 * fN - (float)px, wy = p1 * fN - (float)py, wz = p2 * fN - (float)pz;
  const long long base = ((long long)pz * N + py) * N + px;
  const long long sy = N, sz = (long long)N * N;
  const float c000 = grid[base], c001 = grid[base + 1], c010 = grid[base + sy], c011 = grid[base + sy + 1];
  const float c100 = grid[base + sz], c101 = grid[base + sz + 1], c110 = grid[base + sz + sy],
              c111 = grid[base + sz + sy + 1];
  const float ox = 1.0f - wx, oy = 1.0f - wy, oz = 1.0f - wz;
  float w0 = ox * oy * oz * c000;
  float w1 = wx * oy * oz * c001;
  float w2 = ox * wy * oz * c010;
  float w3 = wx * wy * oz * c011;
  float w4 = ox * oy * wz * c100;
  float w5 = wx * oy * wz * c101;
  float w6 = ox * wy * wz * c110;
  float w7 = wx * wy * wz * c111;
  float res = w0 + w1 + w2 + w3 + w4 + w5 + w6 + w7;
  if (res > 0.2f) return 0.0f;
  return res;
}

// Value and gradient for the LM solver: the value is formed exactly as in
// lookup_f32 (the solver's candidate costs use lookup_f32, so the cost at a
// point does not depend on which of the two evaluated it); the gradient is
// the trilinear interpolant's analytic derivative written compactly, e.g.
// d/dx = N [(c001 - c000) oy oz + (c011 - c010) wy oz + (c101 - c100) oy wz
// + (c111 - c110) wy wz], which is DistanceFloat<Jet<float,3>>'s derivative
// up to FP32 rounding order; +-N per axis outside the box and 0 when the
// value is truncated, as the Jet gives.  About half the registers of
// lookup_jet, whose Jet products keep every zero derivative term.
__device__ inline JetF lookup_value_grad(const float* __restrict__ grid, int N, float p0, float p1, float p2) {
  const float fN = (float)N;
  const int px = trunc_i32f(p0 * fN), py = trunc_i32f(p1 * fN), pz = trunc_i32f(p2 * fN);
  if (px < 0 || py < 0 || pz < 0 || px >= N - 1 || py >= N - 1 || pz >= N - 1) {
    const float edge = (float)(N - 1 - 1e-3);
    float l = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    if (px < 0) { l = l + (-p0) * fN; g0 = -fN; }
    else if (px >= N) { l = l + (p0 * fN - edge); g0 = fN; }
    if (py < 0) { l = l + (-p1) * fN; g1 = -fN; }
    else if (py >= N) { l = l + (p1 * fN - edge); g1 = fN; }
    if (pz < 0) { l = l + (-p2) * fN; g2 = -fN; }
    else if (pz >= N) { l = l + (p2 * fN - edge); g2 = fN; }
    return JetF{l, g0, g1, g2};
  }
  const float wx = p0 * fN - (float)px, wy = p1 * fN - (float)py, wz = p2 * fN - (float)pz;
  const long long base = ((long long)pz * N + py) * N + px;
  const long long sy = N, sz = (long long)N * N;
  const float c000 = grid[base], c001 = grid[base + 1], c010 = grid[base + sy], c011 = grid[base + sy + 1];
  const float c100 = grid[base + sz], c101 = grid[base + sz + 1], c110 = grid[base + sz + sy],
              c111 = grid[base + sz + sy + 1];
  const float ox = 1.0f - wx, oy = 1.0f - wy, oz = 1.0f - wz;
  float w0 = ox * oy * oz * c000;
  float w1 = wx * oy * oz * c001;
  float w2 = ox * wy * oz * c010;
  float w3 = wx * wy * oz * c011;
  float w4 = ox * oy * wz * c100;
  float w5 = wx * oy * wz * c101;
  float w6 = ox * wy * wz * c110;
  float w7 = wx * wy * wz * c111;
  float res = w0 + w1 + w2 + w3 + w4 + w5 + w6 + w7;
  if (res > 0.2f) return JetF{0.0f, 0.0f, 0.0f, 0.0f};
  const float yz00 = oy * oz, yz10 = wy * oz, yz01 = oy * wz, yz11 = wy * wz;
  const float g0 = fN * ((c001 - c000) * yz00 + (c011 - c010) * yz10 + (c101 - c100) * yz01 + (c111 - c110) * yz11);
  const float xz00 = ox * oz, xz10 = wx * oz, xz01 = ox * wz, xz11 = wx * wz;
  const float g1 = fN * ((c010 - c000) * xz00 + (c011 - c001) * xz10 + (c110 - c100) * xz01 + (c111 - c101) * xz11);
  const float xy00 = ox * oy, xy10 = wx * oy, xy01 = ox * wy, xy11 = wx * wy;
  const float g2 = fN * ((c100 - c000) * xy00 + (c101 - c001) * xy10 + (c110 - c010) * xy01 + (c111 - c011) * xy11);
  return JetF{res, g0, g1, g2};
}

// DistanceFloat<Jet<float,3>> seeded with the identity: value and gradient.
__device__ inline JetF lookup_jet(const float* __restrict__ grid, int N, float p0, float p1, float p2) {
  const float fNs = (float)N;
  const int px = trunc_i32f(p0 * fNs), py = trunc_i32f(p1 * fNs), pz = trunc_i32f(p2 * fNs);
  const JetF P0{p0, 1.0f, 0.0f, 0.0f}, P1{p1, 0.0f, 1.0f, 0.0f}, P2{p2, 0.0f, 0.0f, 1.0f};
  const JetF fN = jconst(fNs);
  if (px < 0 || py < 0 || pz < 0 || px >= N - 1 || py >= N - 1 || pz >= N - 1) {
    const JetF edge = jconst((float)(N - 1 - 1e-3));
    JetF l = jconst(0.0f);
    if (px < 0) l = jadd(l, jmul(jneg(P0), fN));
    else if (px >= N) l = jadd(l, jsub(jmul(P0, fN), edge));
    if (py < 0) l = jadd(l, jmul(jneg(P1), fN));
    else if (py >= N) l = jadd(l, jsub(jmul(P1, fN), edge));
    if (pz < 0) l = jadd(l, jmul(jneg(P2), fN));
    else if (pz >= N) l = jadd(l, jsub(jmul(P2, fN), edge));
    return l;
  }
  const JetF one = jconst(1.0f);
  const JetF wx = jsub(jmul(P0, fN), jconst((float)px));
  const JetF wy = jsub(jmul(P1, fN), jconst((float)py));
  const JetF wz = jsub(jmul(P2, fN), jconst((float)pz));
  const long long base = ((long long)pz * N + py) * N + px;
  const long long sy = N, sz = (long long)N * N;
  const JetF c000 = jconst(grid[base]), c001 = jconst(grid[base + 1]);
  const JetF c010 = jconst(grid[base + sy]), c011 = jconst(grid[base + sy + 1]);
  const JetF c100 = jconst(grid[base + sz]), c101 = jconst(grid[base + sz + 1]);
  const JetF c110 = jconst(grid[base + sz + sy]), c111 = jconst(grid[base + sz + sy + 1]);
  const JetF ox = jsub(one, wx), oy = jsub(one, wy), oz = jsub(one, wz);
  JetF w0 = jmul(jmul(jmul(ox, oy), oz), c000);
  JetF w1 = jmul(jmul(jmul(wx, oy), oz), c001);
  JetF w2 = jmul(jmul(jmul(ox, wy), oz), c010);
  JetF w3 = jmul(jmul(jmul(wx, wy), oz), c011);
  JetF w4 = jmul(jmul(jmul(ox, oy), wz), c100);
  JetF w5 = jmul(jmul(jmul(wx, oy), wz), c101);
  JetF w6 = jmul(jmul(jmul(ox, wy), wz), c110);
  JetF w7 = jmul(jmul(jmul(wx, wy), wz), c111);
  JetF res = jadd(jadd(jadd(jadd(jadd(jadd(jadd(w0, w1), w2), w3), w4), w5), w6), w7);
  if (res.a > 0.2f) return jconst(0.0f);
  return res;
}

}  // namespace mo
