// LinearEstimation (linear.cc:11-112) on the GPU.
//
// The reference assembles a sparse SPD system A x = B (three right-hand
// sides) and solves it with Eigen's SimplicialLDLT.  Here the host assembles
// the same matrix (each entry summed in the reference's insertion order) as
// CSR with sorted rows, and one workgroup per system runs Jacobi-
// preconditioned conjugate gradients in FP64 on the three columns in
// lockstep, until every column's residual is below rtol * |b| (a direct
// solve and CG agree to that tolerance; A is SPD: 1e-6 I plus positive
// semidefinite cluster and edge terms).
#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "mo_common.h"
#include "mo_linear.h"
#include "mo_lm_util.h"

namespace mo {
namespace {

constexpr int kCGThreads = 1024;

// one workgroup per system; vectors laid out [n][3]
__global__ __launch_bounds__(kCGThreads) void k_cg3(const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const double* __restrict__ va, const double* __restrict__ dinv,
                                                     const double* __restrict__ B, double* __restrict__ X,
                                                     double* __restrict__ R, double* __restrict__ Z,
                                                     double* __restrict__ P, double* __restrict__ Q, int n,
                                                     int max_iter, double rtol, int* __restrict__ info) {
  __shared__ double red[2 * 8 * kMaxWaves];
  int sel = 0;
  const int tid = threadIdx.x, nt = blockDim.x;
  // r = b - A x, z = D^-1 r, p = z
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int i = tid; i < n; i += nt) {
    double ax[3] = {0, 0, 0};
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const double a = va[e];
      const int j = ci[e];
      for (int c = 0; c < 3; ++c) ax[c] += a * X[3 * j + c];
    }
    for (int c = 0; c < 3; ++c) {
      const double b = B[3 * i + c];
      const double r = b - ax[c];
      const double z = dinv[i] * r;
      R[3 * i + c] = r;
      Z[3 * i + c] = z;
      P[3 * i + c] = z;
      v[c] += r * z;
      v[3 + c] += b * b;
    }
  }
  block_reduce<6, 0>(v, red, sel);
  double rz[3] = {v[0], v[1], v[2]};
  // one tolerance for the three columns, relative to the largest |b|
  const double bmax = fmax(v[3], fmax(v[4], v[5]));
  double tol2[3];
  for (int c = 0; c < 3; ++c) tol2[c] = rtol * rtol * bmax;
  __syncthreads();
  int it = 0;
  bool done = false;
  {
    // initial residual check
    double w[3] = {0, 0, 0};
    for (int i = tid; i < n; i += nt)
      for (int c = 0; c < 3; ++c) w[c] += R[3 * i + c] * R[3 * i + c];
    block_reduce<3, 0>(w, red, sel);
    done = w[0] <= tol2[0] && w[1] <= tol2[1] && w[2] <= tol2[2];
  }
  while (!done && it < max_iter) {
    // q = A p, pq = p . q
    double pq[3] = {0, 0, 0};
    for (int i = tid; i < n; i += nt) {
      double q[3] = {0, 0, 0};
      for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const double a = va[e];
        const int j = ci[e];
        for (int c = 0; c < 3; ++c) q[c] += a * P[3 * j + c];
      }
      for (int c = 0; c < 3; ++c) {
        Q[3 * i + c] = q[c];
        pq[c] += P[3 * i + c] * q[c];
      }
    }
    block_reduce<3, 0>(pq, red, sel);
    double alpha[3];
    for (int c = 0; c < 3; ++c) alpha[c] = pq[c] > 0 ? rz[c] / pq[c] : 0.0;
    // x += alpha p, r -= alpha q, z = D^-1 r; rz' = r . z, rr = r . r
    double w[6] = {0, 0, 0, 0, 0, 0};
    for (int i = tid; i < n; i += nt) {
      for (int c = 0; c < 3; ++c) {
        X[3 * i + c] += alpha[c] * P[3 * i + c];
        const double r = R[3 * i + c] - alpha[c] * Q[3 * i + c];
        const double z = dinv[i] * r;
        R[3 * i + c] = r;
        Z[3 * i + c] = z;
        w[c] += r * z;
        w[3 + c] += r * r;
      }
    }
    block_reduce<6, 0>(w, red, sel);
    ++it;
    done = w[3] <= tol2[0] && w[4] <= tol2[1] && w[5] <= tol2[2];
    if (done) break;
    double beta[3];
    for (int c = 0; c < 3; ++c) {
      beta[c] = rz[c] > 0 ? w[c] / rz[c] : 0.0;
      rz[c] = w[c];
    }
    __syncthreads();  // every thread's Q reads of P are done
    for (int i = tid; i < n; i += nt)
      for (int c = 0; c < 3; ++c) P[3 * i + c] = Z[3 * i + c] + beta[c] * P[3 * i + c];
    __syncthreads();
  }
  if (tid == 0) {
    info[0] = it;
    info[1] = done ? 1 : 0;
  }
}

// A and B of linear.cc:11-105, entries summed in the reference's order.
struct Assembly {
  std::vector<int> rp, ci;
  std::vector<double> va, dinv, B;
};

void assemble(const double* V, int64_t n, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
              const int32_t* ref, const double* GV, double rigidity, Assembly* out) {
  std::unordered_map<long long, size_t> slot;
  std::vector<long long> keys;
  std::vector<double> vals;
  slot.reserve((size_t)n * 16);
  auto addA = [&](int x, int y, double w) {
    const long long k = (long long)x * n + y;
    auto it = slot.find(k);
    if (it == slot.end()) {
      slot.emplace(k, vals.size());
      keys.push_back(k);
      vals.push_back(w);
    } else {
      vals[it->second] += w;
    }
  };
  std::vector<double>& B = out->B;
  B.assign((size_t)n * 3, 0.0);
  int64_t ng = 0;
  for (int64_t i = 0; i < n; ++i) ng = std::max<int64_t>(ng, (int64_t)ref[i] + 1);
  std::vector<std::vector<int>> cell(ng);
  for (int64_t i = 0; i < n; ++i) cell[ref[i]].push_back((int)i);
  for (int64_t g = 0; g < ng; ++g) {
    const double w = 1.0 / cell[g].size();
    const double w2 = w * w;
    for (int a : cell[g])
      for (int b : cell[g]) addA(a, b, w2);
    for (int a : cell[g])
      for (int c = 0; c < 3; ++c) B[3 * a + c] += w * GV[3 * g + c];
  }
  for (int64_t i = 0; i < n; ++i) {
    const double w = 1e-6;
    addA((int)i, (int)i, w);
    for (int c = 0; c < 3; ++c) B[3 * i + c] += w * GV[3 * ref[i] + c];
  }
  auto edge = [&](int a, int b) {
    const double d0 = V[3 * a] - V[3 * b], d1 = V[3 * a + 1] - V[3 * b + 1], d2 = V[3 * a + 2] - V[3 * b + 2];
    double reg = rigidity * 2e-2 / (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) + 1e-8);
    reg *= reg;
    addA(a, a, reg);
    addA(a, b, -reg);
    addA(b, a, -reg);
    addA(b, b, reg);
    for (int c = 0; c < 3; ++c) {
      B[3 * a + c] += reg * (V[3 * a + c] - V[3 * b + c]);
      B[3 * b + c] += reg * (V[3 * b + c] - V[3 * a + c]);
    }
  };
  for (int64_t f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) edge(F[3 * f + j], F[3 * f + (j + 1) % 3]);
  for (int64_t e = 0; e < ne; ++e) edge(E[2 * e], E[2 * e + 1]);
  // CSR with sorted rows
  std::vector<size_t> order(keys.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
  out->rp.assign((size_t)n + 1, 0);
  out->ci.resize(keys.size());
  out->va.resize(keys.size());
  out->dinv.assign((size_t)n, 0.0);
  for (size_t t = 0; t < order.size(); ++t) {
    const long long k = keys[order[t]];
    const int r = (int)(k / n), c = (int)(k % n);
    out->rp[r + 1]++;
    out->ci[t] = c;
    out->va[t] = vals[order[t]];
    if (r == c) out->dinv[r] = 1.0 / vals[order[t]];
  }
  for (int64_t i = 0; i < n; ++i) out->rp[i + 1] += out->rp[i];
}

}  // namespace

int linear_estimation(double* V, int64_t nv, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                      const int32_t* ref, const double* GV, int64_t ngv, double rigidity, int max_iter, double rtol,
                      int* iterations, hipStream_t s) {
  MO_CHECK(nv > 0 && nv < (1LL << 31), MO_ERR_ARG, "linear_estimation: bad vertex count %lld", (long long)nv);
  std::vector<double> hV(nv * 3), hG(ngv * 3);
  std::vector<int32_t> hF(nf * 3), hE(ne * 2), hR(nv);
  MO_HIP(hipMemcpyAsync(hV.data(), V, hV.size() * 8, hipMemcpyDeviceToHost, s));
  if (ngv) MO_HIP(hipMemcpyAsync(hG.data(), GV, hG.size() * 8, hipMemcpyDeviceToHost, s));
  if (nf) MO_HIP(hipMemcpyAsync(hF.data(), F, hF.size() * 4, hipMemcpyDeviceToHost, s));
  if (ne) MO_HIP(hipMemcpyAsync(hE.data(), E, hE.size() * 4, hipMemcpyDeviceToHost, s));
  MO_HIP(hipMemcpyAsync(hR.data(), ref, hR.size() * 4, hipMemcpyDeviceToHost, s));
  MO_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < nv; ++i)
    MO_CHECK(hR[i] >= 0 && hR[i] < ngv, MO_ERR_ARG, "linear_estimation: reference %d of vertex %lld out of range",
             hR[i], (long long)i);
  for (int64_t i = 0; i < nf * 3; ++i)
    MO_CHECK(hF[i] >= 0 && hF[i] < nv, MO_ERR_ARG, "linear_estimation: face index out of range");
  for (int64_t i = 0; i < ne * 2; ++i)
    MO_CHECK(hE[i] >= 0 && hE[i] < nv, MO_ERR_ARG, "linear_estimation: edge index out of range");
  Assembly A;
  assemble(hV.data(), nv, hF.data(), nf, hE.data(), ne, hR.data(), hG.data(), rigidity, &A);
  const size_t nnz = A.va.size();
  const size_t bytes = (nv + 1) * 4 + nnz * 12 + nv * 8 + nv * 24 * 5 + 256 * 8;
  char* ws = nullptr;
  MO_HIP(hipMallocAsync((void**)&ws, bytes, s));
  char* w = ws;
  auto take = [&](size_t b) {
    char* p = w;
    w += (b + 255) & ~size_t(255);
    return p;
  };
  int* d_rp = (int*)take((nv + 1) * 4);
  int* d_ci = (int*)take(nnz * 4);
  double* d_va = (double*)take(nnz * 8);
  double* d_dinv = (double*)take(nv * 8);
  double* d_B = (double*)take(nv * 24);
  double* d_R = (double*)take(nv * 24);
  double* d_Z = (double*)take(nv * 24);
  double* d_P = (double*)take(nv * 24);
  double* d_Q = (double*)take(nv * 24);
  int* d_info = (int*)take(16);
  int rc = MO_OK;
  do {
    if (hipMemcpyAsync(d_rp, A.rp.data(), (nv + 1) * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_ci, A.ci.data(), nnz * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_va, A.va.data(), nnz * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_dinv, A.dinv.data(), nv * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_B, A.B.data(), nv * 24, hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = fail(MO_ERR_HIP, "linear_estimation: upload failed");
      break;
    }
    // the rest positions are the initial guess
    hipLaunchKernelGGL(k_cg3, dim3(1), dim3(kCGThreads), 0, s, d_rp, d_ci, d_va, d_dinv, d_B, V, d_R, d_Z, d_P, d_Q,
                       (int)nv, max_iter, rtol, d_info);
    if (hipGetLastError() != hipSuccess) {
      rc = fail(MO_ERR_HIP, "linear_estimation: launch failed");
      break;
    }
    int info[2] = {0, 0};
    if (hipMemcpyAsync(info, d_info, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(MO_ERR_HIP, "linear_estimation: solve failed");
      break;
    }
    if (iterations) *iterations = info[0];
    if (!info[1]) rc = fail(MO_ERR_NOCONV, "linear_estimation: CG did not reach rtol %g in %d iterations", rtol, info[0]);
  } while (0);
  (void)hipFreeAsync(ws, s);
  return rc;
}

}  // namespace mo
