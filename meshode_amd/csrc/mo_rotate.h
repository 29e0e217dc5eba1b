// Angle-axis rotation of a point with its derivatives, usable on host and
// device: the rotation EdgeLossWithRot applies (edgeloss.h:64-90, through
// Ceres' AngleAxisRotatePoint).  Ceres is absent here; this restates its
// published algorithm: with theta^2 = |w|^2 above double epsilon, Rodrigues'
// formula
//     q = p cos(theta) + (u x p) sin(theta) + u (u . p)(1 - cos(theta)),
//     u = w / theta,
// and below it the first-order form q = p + w x p.  The derivatives are
// those of the same expressions (what Ceres' automatic differentiation
// returns): dq/dp = cos I + sin [u]x + (1 - cos) u u^T (first order:
// I + [w]x), and dq/dw by forward-mode dual numbers through the formula.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mo {

// value and the three partial derivatives with respect to w
struct D3 {
  float a, d0, d1, d2;
};
__host__ __device__ inline D3 d3c(float a) { return D3{a, 0.f, 0.f, 0.f}; }
__host__ __device__ inline D3 operator+(D3 x, D3 y) { return D3{x.a + y.a, x.d0 + y.d0, x.d1 + y.d1, x.d2 + y.d2}; }
__host__ __device__ inline D3 operator-(D3 x, D3 y) { return D3{x.a - y.a, x.d0 - y.d0, x.d1 - y.d1, x.d2 - y.d2}; }
__host__ __device__ inline D3 operator*(D3 x, D3 y) {
  return D3{x.a * y.a, x.d0 * y.a + x.a * y.d0, x.d1 * y.a + x.a * y.d1, x.d2 * y.a + x.a * y.d2};
}
__host__ __device__ inline D3 scale(D3 x, float s) { return D3{x.a * s, x.d0 * s, x.d1 * s, x.d2 * s}; }
__host__ __device__ inline D3 d3sqrt(D3 x) {
  const float r = sqrtf(x.a), h = 0.5f / r;
  return D3{r, x.d0 * h, x.d1 * h, x.d2 * h};
}
__host__ __device__ inline D3 d3inv(D3 x) {
  const float r = 1.0f / x.a, h = -r * r;
  return D3{r, x.d0 * h, x.d1 * h, x.d2 * h};
}
__host__ __device__ inline D3 d3sin(D3 x) {
  const float s = sinf(x.a), c = cosf(x.a);
  return D3{s, x.d0 * c, x.d1 * c, x.d2 * c};
}
__host__ __device__ inline D3 d3cos(D3 x) {
  const float s = sinf(x.a), c = cosf(x.a);
  return D3{c, -x.d0 * s, -x.d1 * s, -x.d2 * s};
}

constexpr float kAngleAxisEps = 2.220446049250313e-16f;   // double epsilon, the branch threshold

// dq/dp (row-major 3x3) at w: depends on w only.
__host__ __device__ inline void rotation_point_jacobian(const float w[3], float A[9]) {
  const float t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (t2 > kAngleAxisEps) {
    const float t = sqrtf(t2), c = cosf(t), s = sinf(t), it = 1.0f / t;
    const float u[3] = {w[0] * it, w[1] * it, w[2] * it};
    const float oc = 1.0f - c;
    A[0] = c + oc * u[0] * u[0];       A[1] = -s * u[2] + oc * u[0] * u[1]; A[2] = s * u[1] + oc * u[0] * u[2];
    A[3] = s * u[2] + oc * u[1] * u[0]; A[4] = c + oc * u[1] * u[1];      A[5] = -s * u[0] + oc * u[1] * u[2];
    A[6] = -s * u[1] + oc * u[2] * u[0]; A[7] = s * u[0] + oc * u[2] * u[1]; A[8] = c + oc * u[2] * u[2];
  } else {
    A[0] = 1.f;   A[1] = -w[2]; A[2] = w[1];
    A[3] = w[2];  A[4] = 1.f;   A[5] = -w[0];
    A[6] = -w[1]; A[7] = w[0];  A[8] = 1.f;
  }
}

// q = rotate(w, p) and G = dq/dw (row-major: G[3 r + c] = dq_r / dw_c).
__host__ __device__ inline void rotate_point_dw(const float w[3], const float p[3], float q[3], float G[9]) {
  const D3 W[3] = {D3{w[0], 1.f, 0.f, 0.f}, D3{w[1], 0.f, 1.f, 0.f}, D3{w[2], 0.f, 0.f, 1.f}};
  const D3 P[3] = {d3c(p[0]), d3c(p[1]), d3c(p[2])};
  const D3 t2 = W[0] * W[0] + W[1] * W[1] + W[2] * W[2];
  D3 Q[3];
  if (t2.a > kAngleAxisEps) {
    const D3 t = d3sqrt(t2);
    const D3 c = d3cos(t), s = d3sin(t), it = d3inv(t);
    const D3 u[3] = {W[0] * it, W[1] * it, W[2] * it};
    const D3 uxp[3] = {u[1] * P[2] - u[2] * P[1], u[2] * P[0] - u[0] * P[2], u[0] * P[1] - u[1] * P[0]};
    const D3 tmp = (u[0] * P[0] + u[1] * P[1] + u[2] * P[2]) * (d3c(1.0f) - c);
    for (int k = 0; k < 3; ++k) Q[k] = P[k] * c + uxp[k] * s + u[k] * tmp;
  } else {
    const D3 wxp[3] = {W[1] * P[2] - W[2] * P[1], W[2] * P[0] - W[0] * P[2], W[0] * P[1] - W[1] * P[0]};
    for (int k = 0; k < 3; ++k) Q[k] = P[k] + wxp[k];
  }
  for (int k = 0; k < 3; ++k) {
    q[k] = Q[k].a;
    G[3 * k] = Q[k].d0;
    G[3 * k + 1] = Q[k].d1;
    G[3 * k + 2] = Q[k].d2;
  }
}

// q = rotate(w, p) only.
__host__ __device__ inline void rotate_point(const float w[3], const float p[3], float q[3]) {
  const float t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (t2 > kAngleAxisEps) {
    const float t = sqrtf(t2), c = cosf(t), s = sinf(t), it = 1.0f / t;
    const float u[3] = {w[0] * it, w[1] * it, w[2] * it};
    const float uxp[3] = {u[1] * p[2] - u[2] * p[1], u[2] * p[0] - u[0] * p[2], u[0] * p[1] - u[1] * p[0]};
    const float tmp = (u[0] * p[0] + u[1] * p[1] + u[2] * p[2]) * (1.0f - c);
    for (int k = 0; k < 3; ++k) q[k] = p[k] * c + uxp[k] * s + u[k] * tmp;
  } else {
    q[0] = p[0] + (w[1] * p[2] - w[2] * p[1]);
    q[1] = p[1] + (w[2] * p[0] - w[0] * p[2]);
    q[2] = p[2] + (w[0] * p[1] - w[1] * p[0]);
  }
}

}  // namespace mo
