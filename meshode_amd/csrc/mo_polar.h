// Orthogonal polar factor of a 3x3 matrix, usable on host and device.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mo {

// R = U V^T of the SVD of a 3x3 matrix C (row-major): the orthogonal polar
// factor, a reflection when det C < 0 (LinearEstimationWithRot takes it of
// each vertex's neighbourhood covariance, linear.cc:158-163).
// V comes from cyclic Jacobi on C^T C; u_k = C v_k / |C v_k| with u1
// orthogonalized against u0 and u2 = +-(u0 x u1), the sign of C v2.  Where
// C is singular the product R (V_j - V_i) does not depend on the choice of
// the null-space vectors, since (V_j - V_i) lies in the row space.
__host__ __device__ inline void polar3(const double* C, double* R) {
  double M[3][3], Q[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      M[i][j] = C[i] * C[j] + C[3 + i] * C[3 + j] + C[6 + i] * C[6 + j];
      Q[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 24; ++sweep) {
    const double off = M[0][1] * M[0][1] + M[0][2] * M[0][2] + M[1][2] * M[1][2];
    const double dia = M[0][0] * M[0][0] + M[1][1] * M[1][1] + M[2][2] * M[2][2];
    if (off <= 1e-36 * dia) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      const double apq = M[p][q];
      if (apq == 0.0) continue;
      const double theta = (M[q][q] - M[p][p]) / (2.0 * apq);
      const double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      for (int k = 0; k < 3; ++k) {
        const double x = M[k][p], y = M[k][q];
        M[k][p] = c * x - sn * y;
        M[k][q] = sn * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = M[p][k], y = M[q][k];
        M[p][k] = c * x - sn * y;
        M[q][k] = sn * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = Q[k][p], y = Q[k][q];
        Q[k][p] = c * x - sn * y;
        Q[k][q] = sn * x + c * y;
      }
    }
  }
  int o[3] = {0, 1, 2};
  if (M[o[1]][o[1]] > M[o[0]][o[0]]) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
  if (M[o[2]][o[2]] > M[o[1]][o[1]]) { const int t = o[1]; o[1] = o[2]; o[2] = t; }
  if (M[o[1]][o[1]] > M[o[0]][o[0]]) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
  double v[3][3], u[3][3];
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) v[k][r] = Q[r][o[k]];
  auto mulc = [&](const double* x, double* y) {
    for (int r = 0; r < 3; ++r) y[r] = C[3 * r] * x[0] + C[3 * r + 1] * x[1] + C[3 * r + 2] * x[2];
  };
  mulc(v[0], u[0]);
  const double n0 = sqrt(u[0][0] * u[0][0] + u[0][1] * u[0][1] + u[0][2] * u[0][2]);
  if (!(n0 > 0.0)) {  // zero covariance: identity, as an SVD of 0 gives
    for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  for (int r = 0; r < 3; ++r) u[0][r] /= n0;
  mulc(v[1], u[1]);
  const double d01 = u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2];
  for (int r = 0; r < 3; ++r) u[1][r] -= d01 * u[0][r];
  double n1 = sqrt(u[1][0] * u[1][0] + u[1][1] * u[1][1] + u[1][2] * u[1][2]);
  if (!(n1 > 0.0)) {  // rank one: any unit vector orthogonal to u0
    const int ax = fabs(u[0][0]) <= fabs(u[0][1]) ? (fabs(u[0][0]) <= fabs(u[0][2]) ? 0 : 2)
                                                  : (fabs(u[0][1]) <= fabs(u[0][2]) ? 1 : 2);
    double e[3] = {0, 0, 0};
    e[ax] = 1.0;
    u[1][0] = e[1] * u[0][2] - e[2] * u[0][1];
    u[1][1] = e[2] * u[0][0] - e[0] * u[0][2];
    u[1][2] = e[0] * u[0][1] - e[1] * u[0][0];
    n1 = sqrt(u[1][0] * u[1][0] + u[1][1] * u[1][1] + u[1][2] * u[1][2]);
  }
  for (int r = 0; r < 3; ++r) u[1][r] /= n1;
  u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
  u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
  u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
  double cv2[3];
  mulc(v[2], cv2);
  if (cv2[0] * u[2][0] + cv2[1] * u[2][1] + cv2[2] * u[2][2] < 0.0)
    for (int r = 0; r < 3; ++r) u[2][r] = -u[2][r];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[3 * r + c] = u[0][r] * v[0][c] + u[1][r] * v[1][c] + u[2][r] * v[2][c];
}

}  // namespace mo
