// LinearEstimation (linear.cc:11-112) on the GPU.
//
// The reference minimizes the quadratic
//   sum_g |mean_{v in g} x_v - GV_g|^2 + 1e-6 sum_v |x_v - GV_ref(v)|^2
//   + sum_{edges} w_e |(x_a - x_b) - (V_a - V_b)|^2,
//   w_e = (rigidity 2e-2 / (|V_a - V_b| + 1e-8))^2,
// over every face half-edge and every geometry-neighbour pair, by
// assembling A x = B and factoring A with Eigen's SimplicialLDLT.  The
// subdivision creates near-coincident vertices (lattice points next to edge
// points), whose edges get weights up to 1.6e13; an assembled FP64 matrix
// then carries the soft terms of those rows with an absolute error of
// eps * 1.6e13 ~ 4e-3, and a direct solve of it is off by ~1e-4.
//
// Here the same energy is minimized with products that never form the
// assembled rows: A p = 1e-6 p + sum_j w_ij (p_i - p_j) + w_g^2 sum_{k in g} p_k,
// so the equal and opposite edge forces cancel exactly in the common motion
// of a stiff pair.  One workgroup runs preconditioned conjugate gradients in
// FP64 on the three columns in lockstep; the preconditioner inverts exactly
// the blocks of vertices joined by stiff edges (weight > 1e4 x the median)
// and the diagonal elsewhere.  B is assembled on the host in the
// reference's order (linear.cc:38-89).
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "mo_common.h"
#include "mo_linear.h"
#include "mo_lm_util.h"

namespace mo {
namespace {

constexpr int kCGThreads = 1024;
constexpr int kMaxBlock = 64;

struct CGArgs {
  int n, ng;
  const int* rp;        // [n+1] Laplacian rows
  const int* ci;        // neighbours
  const double* w;      // merged edge weights
  const int* cid;       // [n] cluster of each vertex
  const int* cp;        // [ng+1] cluster members
  const int* cm;
  const double* cw2;    // [ng] squared cluster weight 1/|g|^2
  const int* pstart;    // [n] preconditioner block of each vertex: members bmem[pstart .. pstart+psize)
  const int* psize;
  const int* prow;      // [n] this vertex's row of the block inverse in binv
  const int* bmem;
  const double* binv;
  const double* B;      // [n,3]
  double* X;            // [n,3] in: rest positions, out: solution
  double* R;
  double* Z;
  double* P;
  double* Q;
  double* S;            // [ng,3] cluster sums
  int max_iter;
  double rtol;
  int* info;
};

__device__ inline void cluster_sums(const CGArgs& a, const double* __restrict__ Y) {
  for (int g = threadIdx.x; g < a.ng; g += blockDim.x) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = a.cp[g]; k < a.cp[g + 1]; ++k) {
      const int m = a.cm[k];
      s0 += Y[3 * m];
      s1 += Y[3 * m + 1];
      s2 += Y[3 * m + 2];
    }
    a.S[3 * g] = s0;
    a.S[3 * g + 1] = s1;
    a.S[3 * g + 2] = s2;
  }
}

// (A y)_i from edge differences, the cluster sums and the 1e-6 diagonal
__device__ inline void apply_row(const CGArgs& a, const double* __restrict__ Y, int i, double* out) {
  const double yi[3] = {Y[3 * i], Y[3 * i + 1], Y[3 * i + 2]};
  double acc[3] = {0, 0, 0};
  for (int e = a.rp[i]; e < a.rp[i + 1]; ++e) {
    const int j = a.ci[e];
    const double we = a.w[e];
    for (int c = 0; c < 3; ++c) acc[c] += we * (yi[c] - Y[3 * j + c]);
  }
  const int g = a.cid[i];
  for (int c = 0; c < 3; ++c) out[c] = 1e-6 * yi[c] + acc[c] + a.cw2[g] * a.S[3 * g + c];
}

__device__ inline void precond_row(const CGArgs& a, const double* __restrict__ Rv, int i, double* out) {
  const int s = a.pstart[i], m = a.psize[i], row = a.prow[i];
  double z[3] = {0, 0, 0};
  for (int k = 0; k < m; ++k) {
    const double b = a.binv[row + k];
    const int j = a.bmem[s + k];
    for (int c = 0; c < 3; ++c) z[c] += b * Rv[3 * j + c];
  }
  for (int c = 0; c < 3; ++c) out[c] = z[c];
}

__global__ __launch_bounds__(kCGThreads) void k_cg3(CGArgs a) {
  __shared__ double red[2 * 8 * kMaxWaves];
  int sel = 0;
  const int tid = threadIdx.x, nt = blockDim.x, n = a.n;
  // r = b - A x
  cluster_sums(a, a.X);
  __syncthreads();
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = tid; i < n; i += nt) {
    double ax[3];
    apply_row(a, a.X, i, ax);
    for (int c = 0; c < 3; ++c) {
      const double b = a.B[3 * i + c];
      const double r = b - ax[c];
      a.R[3 * i + c] = r;
      v[3 + c] += b * b;
      v[6] = fmax(v[6], fabs(r));
    }
  }
  __syncthreads();
  // z = M^-1 r, p = z
  for (int i = tid; i < n; i += nt) {
    double z[3];
    precond_row(a, a.R, i, z);
    for (int c = 0; c < 3; ++c) {
      a.Z[3 * i + c] = z[c];
      a.P[3 * i + c] = z[c];
      v[c] += a.R[3 * i + c] * z[c];
    }
  }
  block_reduce<6, 1>(v, red, sel);
  double rz[3] = {v[0], v[1], v[2]};
  // one tolerance for the three columns, relative to the largest |b|
  const double tol2 = a.rtol * a.rtol * fmax(v[3], fmax(v[4], v[5]));
  bool done = v[6] == 0.0;
  int it = 0;
  while (!done && it < a.max_iter) {
    cluster_sums(a, a.P);
    __syncthreads();
    double pq[3] = {0, 0, 0};
    for (int i = tid; i < n; i += nt) {
      double q[3];
      apply_row(a, a.P, i, q);
      for (int c = 0; c < 3; ++c) {
        a.Q[3 * i + c] = q[c];
        pq[c] += a.P[3 * i + c] * q[c];
      }
    }
    block_reduce<3, 0>(pq, red, sel);
    double alpha[3];
    for (int c = 0; c < 3; ++c) alpha[c] = pq[c] > 0 ? rz[c] / pq[c] : 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    for (int i = tid; i < n; i += nt)
      for (int c = 0; c < 3; ++c) {
        a.X[3 * i + c] += alpha[c] * a.P[3 * i + c];
        const double r = a.R[3 * i + c] - alpha[c] * a.Q[3 * i + c];
        a.R[3 * i + c] = r;
        w[3 + c] += r * r;
      }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
      double z[3];
      precond_row(a, a.R, i, z);
      for (int c = 0; c < 3; ++c) {
        a.Z[3 * i + c] = z[c];
        w[c] += a.R[3 * i + c] * z[c];
      }
    }
    block_reduce<6, 0>(w, red, sel);
    ++it;
    done = w[3] <= tol2 && w[4] <= tol2 && w[5] <= tol2;
    if (done) break;
    double beta[3];
    for (int c = 0; c < 3; ++c) {
      beta[c] = rz[c] > 0 ? w[c] / rz[c] : 0.0;
      rz[c] = w[c];
    }
    for (int i = tid; i < n; i += nt)
      for (int c = 0; c < 3; ++c) a.P[3 * i + c] = a.Z[3 * i + c] + beta[c] * a.P[3 * i + c];
    __syncthreads();
  }
  if (tid == 0) {
    a.info[0] = it;
    a.info[1] = done ? 1 : 0;
  }
}

struct System {
  int n = 0, ng = 0;
  std::vector<int> rp, ci, cid, cp, cm, pstart, psize, prow, bmem;
  std::vector<double> w, cw2, binv, B;
};

// Gauss-Jordan inverse with partial pivoting (long double), m x m row-major.
void invert(std::vector<long double>& M, int m, double* out) {
  std::vector<long double> I((size_t)m * m, 0.0L);
  for (int i = 0; i < m; ++i) I[(size_t)i * m + i] = 1.0L;
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(M[(size_t)r * m + c]) > std::fabs(M[(size_t)p * m + c])) p = r;
    if (p != c)
      for (int k = 0; k < m; ++k) {
        std::swap(M[(size_t)p * m + k], M[(size_t)c * m + k]);
        std::swap(I[(size_t)p * m + k], I[(size_t)c * m + k]);
      }
    const long double d = M[(size_t)c * m + c];
    for (int k = 0; k < m; ++k) {
      M[(size_t)c * m + k] /= d;
      I[(size_t)c * m + k] /= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const long double f = M[(size_t)r * m + c];
      if (f == 0) continue;
      for (int k = 0; k < m; ++k) {
        M[(size_t)r * m + k] -= f * M[(size_t)c * m + k];
        I[(size_t)r * m + k] -= f * I[(size_t)c * m + k];
      }
    }
  }
  for (size_t i = 0; i < I.size(); ++i) out[i] = (double)I[i];
}

void build_system(const double* V, int64_t n, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                  const int32_t* ref, const double* GV, double rigidity, System* S) {
  S->n = (int)n;
  int ng = 0;
  for (int64_t i = 0; i < n; ++i) ng = std::max(ng, ref[i] + 1);
  S->ng = ng;
  // clusters (grid_cell, linear.cc:32-34): members in vertex order
  S->cid.assign(ref, ref + n);
  S->cp.assign(ng + 1, 0);
  for (int64_t i = 0; i < n; ++i) S->cp[ref[i] + 1]++;
  for (int g = 0; g < ng; ++g) S->cp[g + 1] += S->cp[g];
  S->cm.resize(n);
  {
    std::vector<int> fill(S->cp.begin(), S->cp.end() - 1);
    for (int64_t i = 0; i < n; ++i) S->cm[fill[ref[i]]++] = (int)i;
  }
  S->cw2.assign(ng, 0.0);
  // right-hand side in the reference's order (linear.cc:38-89)
  std::vector<double>& B = S->B;
  B.assign((size_t)n * 3, 0.0);
  for (int g = 0; g < ng; ++g) {
    const int sz = S->cp[g + 1] - S->cp[g];
    if (!sz) continue;
    const double wt = 1.0 / sz;
    S->cw2[g] = wt * wt;
    for (int k = S->cp[g]; k < S->cp[g + 1]; ++k)
      for (int c = 0; c < 3; ++c) B[3 * S->cm[k] + c] += wt * GV[3 * g + c];
  }
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) B[3 * i + c] += 1e-6 * GV[3 * ref[i] + c];
  std::unordered_map<uint64_t, double> wsum;
  wsum.reserve((size_t)(3 * nf + ne) * 2);
  auto edge = [&](int a, int b) {
    const double d0 = V[3 * a] - V[3 * b], d1 = V[3 * a + 1] - V[3 * b + 1], d2 = V[3 * a + 2] - V[3 * b + 2];
    double reg = rigidity * 2e-2 / (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) + 1e-8);
    reg *= reg;
    for (int c = 0; c < 3; ++c) {
      B[3 * a + c] += reg * (V[3 * a + c] - V[3 * b + c]);
      B[3 * b + c] += reg * (V[3 * b + c] - V[3 * a + c]);
    }
    if (a == b) return;  // contributes nothing to A (reg - reg)
    const uint64_t k = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
    wsum[k] += reg;
  };
  for (int64_t f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) edge(F[3 * f + j], F[3 * f + (j + 1) % 3]);
  for (int64_t e = 0; e < ne; ++e) edge(E[2 * e], E[2 * e + 1]);
  // Laplacian rows, neighbours sorted
  std::vector<std::pair<uint64_t, double>> ed(wsum.begin(), wsum.end());
  std::sort(ed.begin(), ed.end());
  S->rp.assign(n + 1, 0);
  for (const auto& e : ed) {
    S->rp[(e.first >> 32) + 1]++;
    S->rp[(uint32_t)e.first + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) S->rp[i + 1] += S->rp[i];
  S->ci.resize(S->rp[n]);
  S->w.resize(S->rp[n]);
  {
    std::vector<int> fill(S->rp.begin(), S->rp.end() - 1);
    std::vector<std::pair<int, double>> tmp;
    for (const auto& e : ed) {
      const int a = (int)(e.first >> 32), b = (int)(uint32_t)e.first;
      S->ci[fill[a]] = b;
      S->w[fill[a]++] = e.second;
      S->ci[fill[b]] = a;
      S->w[fill[b]++] = e.second;
    }
    for (int64_t i = 0; i < n; ++i) {
      tmp.clear();
      for (int e = S->rp[i]; e < S->rp[i + 1]; ++e) tmp.emplace_back(S->ci[e], S->w[e]);
      std::sort(tmp.begin(), tmp.end());
      for (size_t k = 0; k < tmp.size(); ++k) {
        S->ci[S->rp[i] + k] = tmp[k].first;
        S->w[S->rp[i] + k] = tmp[k].second;
      }
    }
  }
  // preconditioner blocks: vertices joined by stiff edges (union-find,
  // at most kMaxBlock vertices per block)
  std::vector<double> ws;
  for (const auto& e : ed) ws.push_back(e.second);
  double tau = INFINITY;
  if (!ws.empty()) {
    std::nth_element(ws.begin(), ws.begin() + ws.size() / 2, ws.end());
    tau = 1e4 * ws[ws.size() / 2];
  }
  std::vector<int> parent(n), size(n, 1);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (const auto& e : ed) {
    if (!(e.second > tau)) continue;
    int a = find((int)(e.first >> 32)), b = find((int)(uint32_t)e.first);
    if (a == b || size[a] + size[b] > kMaxBlock) continue;
    if (size[a] < size[b]) std::swap(a, b);
    parent[b] = a;
    size[a] += size[b];
  }
  std::vector<std::vector<int>> members(n);
  for (int64_t i = 0; i < n; ++i) members[find((int)i)].push_back((int)i);
  S->pstart.assign(n, 0);
  S->psize.assign(n, 0);
  S->prow.assign(n, 0);
  auto diag = [&](int i) {
    double d = 1e-6 + S->cw2[S->cid[i]];
    for (int e = S->rp[i]; e < S->rp[i + 1]; ++e) d += S->w[e];
    return d;
  };
  std::vector<long double> M;
  for (int64_t r = 0; r < n; ++r) {
    const std::vector<int>& mb = members[r];
    if (mb.empty()) continue;
    const int m = (int)mb.size();
    const int start = (int)S->bmem.size();
    S->bmem.insert(S->bmem.end(), mb.begin(), mb.end());
    const size_t row0 = S->binv.size();
    S->binv.resize(row0 + (size_t)m * m);
    if (m == 1) {
      S->binv[row0] = 1.0 / diag(mb[0]);
    } else {
      M.assign((size_t)m * m, 0.0L);
      for (int x = 0; x < m; ++x) {
        const int i = mb[x];
        M[(size_t)x * m + x] = diag(i);
        for (int y = 0; y < m; ++y) {
          const int j = mb[y];
          if (y != x && S->cid[i] == S->cid[j]) M[(size_t)x * m + y] += S->cw2[S->cid[i]];
        }
        for (int e = S->rp[i]; e < S->rp[i + 1]; ++e) {
          const int j = S->ci[e];
          for (int y = 0; y < m; ++y)
            if (mb[y] == j) M[(size_t)x * m + y] -= S->w[e];
        }
      }
      invert(M, m, &S->binv[row0]);
    }
    for (int x = 0; x < m; ++x) {
      S->pstart[mb[x]] = start;
      S->psize[mb[x]] = m;
      S->prow[mb[x]] = (int)(row0 + (size_t)x * m);
    }
  }
}

}  // namespace

int linear_estimation(double* V, int64_t nv, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                      const int32_t* ref, const double* GV, int64_t ngv, double rigidity, int max_iter, double rtol,
                      int* iterations, hipStream_t s) {
  MO_CHECK(nv > 0 && nv < (1LL << 30), MO_ERR_ARG, "linear_estimation: bad vertex count %lld", (long long)nv);
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
  System S;
  build_system(hV.data(), nv, hF.data(), nf, hE.data(), ne, hR.data(), hG.data(), rigidity, &S);
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t nn = (size_t)nv, ng = (size_t)S.ng, nnz = S.ci.size(), nb = S.bmem.size(), ni = S.binv.size();
  const size_t bytes = al((nn + 1) * 4) + al(nnz * 4) + al(nnz * 8) + al(nn * 4) + al((ng + 1) * 4) + al(nn * 4) +
                       al(ng * 8) + 3 * al(nn * 4) + al(nb * 4) + al(ni * 8) + 5 * al(nn * 24) + al(ng * 24) + al(16);
  char* ws = nullptr;
  MO_HIP(hipMallocAsync((void**)&ws, bytes, s));
  char* w = ws;
  auto take = [&](size_t b) {
    char* p = w;
    w += al(b);
    return p;
  };
  CGArgs a;
  a.n = (int)nv;
  a.ng = S.ng;
  int* d_rp = (int*)take((nn + 1) * 4);
  int* d_ci = (int*)take(nnz * 4);
  double* d_w = (double*)take(nnz * 8);
  int* d_cid = (int*)take(nn * 4);
  int* d_cp = (int*)take((ng + 1) * 4);
  int* d_cm = (int*)take(nn * 4);
  double* d_cw2 = (double*)take(ng * 8);
  int* d_ps = (int*)take(nn * 4);
  int* d_pz = (int*)take(nn * 4);
  int* d_pr = (int*)take(nn * 4);
  int* d_bm = (int*)take(nb * 4);
  double* d_bi = (double*)take(ni * 8);
  double* d_B = (double*)take(nn * 24);
  a.R = (double*)take(nn * 24);
  a.Z = (double*)take(nn * 24);
  a.P = (double*)take(nn * 24);
  a.Q = (double*)take(nn * 24);
  a.S = (double*)take(ng * 24);
  a.info = (int*)take(16);
  a.rp = d_rp; a.ci = d_ci; a.w = d_w; a.cid = d_cid; a.cp = d_cp; a.cm = d_cm; a.cw2 = d_cw2;
  a.pstart = d_ps; a.psize = d_pz; a.prow = d_pr; a.bmem = d_bm; a.binv = d_bi; a.B = d_B;
  a.X = V;
  a.max_iter = max_iter;
  a.rtol = rtol;
  int rc = MO_OK;
  do {
    struct Up { void* d; const void* h; size_t b; } ups[] = {
        {d_rp, S.rp.data(), (nn + 1) * 4}, {d_ci, S.ci.data(), nnz * 4}, {d_w, S.w.data(), nnz * 8},
        {d_cid, S.cid.data(), nn * 4}, {d_cp, S.cp.data(), (ng + 1) * 4}, {d_cm, S.cm.data(), nn * 4},
        {d_cw2, S.cw2.data(), ng * 8}, {d_ps, S.pstart.data(), nn * 4}, {d_pz, S.psize.data(), nn * 4},
        {d_pr, S.prow.data(), nn * 4}, {d_bm, S.bmem.data(), nb * 4}, {d_bi, S.binv.data(), ni * 8},
        {d_B, S.B.data(), nn * 24}};
    bool ok = true;
    for (const Up& u : ups)
      if (u.b && hipMemcpyAsync(u.d, u.h, u.b, hipMemcpyHostToDevice, s) != hipSuccess) ok = false;
    if (!ok) {
      rc = fail(MO_ERR_HIP, "linear_estimation: upload failed");
      break;
    }
    // the rest positions are the initial guess
    hipLaunchKernelGGL(k_cg3, dim3(1), dim3(kCGThreads), 0, s, a);
    if (hipGetLastError() != hipSuccess) {
      rc = fail(MO_ERR_HIP, "linear_estimation: launch failed");
      break;
    }
    int info[2] = {0, 0};
    if (hipMemcpyAsync(info, a.info, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
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
