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
//
// LinearEstimationWithRot (linear.cc:114-230) minimizes
//   sum_i |x_i - TV_i|^2 + sum_{face half-edges (a,b)} sum_{(v0,v1) in {(a,b),(b,a)}}
//       r_ab |(x_v1 - x_v0) - s_v0 R_v0 (V_v1 - V_v0)|^2,   r_ab = 2e-2 rigidity / |V_b - V_a|,
// with s_i R_i the scaled polar factor of vertex i's neighbourhood
// covariance sum_j (TV_j - TV_i)(V_j - V_i)^T.  Per undirected pair (i,j)
// joined by m face half-edges this is the Laplacian weight w_ij = 2 m r_ij
// and the right-hand side TV_i + sum_j (w_ij / 2)(S_i + S_j)(V_i - V_j);
// frames and right-hand side are built on the device (one thread per
// vertex), and the same CG kernel solves with unit diagonal and no
// clusters.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <unordered_map>
#include <vector>

#include <cstdlib>
#include <mutex>
#include "mo_common.h"
#include "mo_linear.h"
#include "mo_polar.h"
#include "mo_lm.h"
#include "mo_lm_util.h"

namespace mo {
namespace {

constexpr int kCGThreads = 1024;
constexpr int kTeamRows = 4096;   // systems this large run on a team of workgroups
constexpr int kMaxBlock = 64;

struct CGArgs {
  int n, ng;            // ng = 0: no cluster term
  double c0;            // diagonal coefficient (1e-6 LinearEstimation, 1 WithRot)
  const double* dg;     // [n] per-row diagonal coefficients, or null for c0
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

__device__ inline void cluster_sums(const CGArgs& a, const double* __restrict__ Y, int g0, int gstride) {
  for (int g = g0; g < a.ng; g += gstride) {
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

// (A y)_i from edge differences, the cluster sums and the c0 diagonal
__device__ inline void apply_row(const CGArgs& a, const double* __restrict__ Y, int i, double* out) {
  const double yi[3] = {Y[3 * i], Y[3 * i + 1], Y[3 * i + 2]};
  double acc[3] = {0, 0, 0};
  for (int e = a.rp[i]; e < a.rp[i + 1]; ++e) {
    const int j = a.ci[e];
    const double we = a.w[e];
    for (int c = 0; c < 3; ++c) acc[c] += we * (yi[c] - Y[3 * j + c]);
  }
  double cl[3] = {0, 0, 0};
  if (a.ng) {
    const int g = a.cid[i];
    for (int c = 0; c < 3; ++c) cl[c] = a.cw2[g] * a.S[3 * g + c];
  }
  const double d0 = a.dg ? a.dg[i] : a.c0;
  for (int c = 0; c < 3; ++c) out[c] = d0 * yi[c] + acc[c] + cl[c];
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

// Preconditioned CG on the three right-hand sides, by one workgroup
// (SoloTeam) or by all workgroups of a cooperative grid (GridTeam: rows
// split as in the LM team solver, FP64 partials summed in rank order).
template <class Team>
__device__ inline void cg3_body(const CGArgs& a, Team& team, double* red) {
  int sel = 0;
  const int tid = threadIdx.x, n = a.n;
  const int i0 = team.rank() * blockDim.x + tid, nt = team.size() * blockDim.x;
  // r = b - A x
  cluster_sums(a, a.X, i0, nt);
  team.barrier();
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = i0; i < n; i += nt) {
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
  team.barrier();
  // z = M^-1 r, p = z
  for (int i = i0; i < n; i += nt) {
    double z[3];
    precond_row(a, a.R, i, z);
    for (int c = 0; c < 3; ++c) {
      a.Z[3 * i + c] = z[c];
      a.P[3 * i + c] = z[c];
      v[c] += a.R[3 * i + c] * z[c];
    }
  }
  team.template reduce<6, 1>(v, red, sel);
  double rz[3] = {v[0], v[1], v[2]};
  // one tolerance for the three columns, relative to the largest |b|
  const double tol2 = a.rtol * a.rtol * fmax(v[3], fmax(v[4], v[5]));
  bool done = v[6] == 0.0;
  int it = 0;
  while (!done && it < a.max_iter) {
    cluster_sums(a, a.P, i0, nt);
    team.barrier();
    double pq[3] = {0, 0, 0};
    for (int i = i0; i < n; i += nt) {
      double q[3];
      apply_row(a, a.P, i, q);
      for (int c = 0; c < 3; ++c) {
        a.Q[3 * i + c] = q[c];
        pq[c] += a.P[3 * i + c] * q[c];
      }
    }
    team.template reduce<3, 0>(pq, red, sel);
    double alpha[3];
    for (int c = 0; c < 3; ++c) alpha[c] = pq[c] > 0 ? rz[c] / pq[c] : 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    for (int i = i0; i < n; i += nt)
      for (int c = 0; c < 3; ++c) {
        a.X[3 * i + c] += alpha[c] * a.P[3 * i + c];
        const double r = a.R[3 * i + c] - alpha[c] * a.Q[3 * i + c];
        a.R[3 * i + c] = r;
        w[3 + c] += r * r;
      }
    team.barrier();
    for (int i = i0; i < n; i += nt) {
      double z[3];
      precond_row(a, a.R, i, z);
      for (int c = 0; c < 3; ++c) {
        a.Z[3 * i + c] = z[c];
        w[c] += a.R[3 * i + c] * z[c];
      }
    }
    team.template reduce<6, 0>(w, red, sel);
    ++it;
    done = w[3] <= tol2 && w[4] <= tol2 && w[5] <= tol2;
    if (done) break;
    double beta[3];
    for (int c = 0; c < 3; ++c) {
      beta[c] = rz[c] > 0 ? w[c] / rz[c] : 0.0;
      rz[c] = w[c];
    }
    for (int i = i0; i < n; i += nt)
      for (int c = 0; c < 3; ++c) a.P[3 * i + c] = a.Z[3 * i + c] + beta[c] * a.P[3 * i + c];
    team.barrier();
  }
  if (tid == 0 && team.leader()) {
    a.info[0] = it;
    a.info[1] = done ? 1 : 0;
  }
}

__global__ void k_zero_words(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kCGThreads) void k_cg3(CGArgs a) {
  __shared__ double red[2 * 8 * kMaxWaves];
  SoloTeam team;
  cg3_body(a, team, red);
}

__global__ __launch_bounds__(kCGThreads) void k_cg3_team(CGArgs a, unsigned* cnt, double* part, int* abort_flag,
                                                          long long limit) {
  __shared__ double red[2 * 8 * kMaxWaves];
  GridTeam team{(int)blockIdx.x, (int)gridDim.x, cnt, part, abort_flag, limit, 0u, 0};
  cg3_body(a, team, red);
}

struct System {
  int n = 0, ng = 0;
  double c0 = 1e-6;
  std::vector<double> dg;  // per-row diagonal coefficients (empty: c0)
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

void finish_system(const std::unordered_map<uint64_t, double>& wsum, System* S);

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
  finish_system(wsum, S);
}

// Laplacian rows (neighbours sorted) from merged pair weights keyed
// (min << 32 | max), and the preconditioner blocks.
void finish_system(const std::unordered_map<uint64_t, double>& wsum, System* S) {
  const int64_t n = S->n;
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
    double d = (S->dg.empty() ? S->c0 : S->dg[i]) + (S->ng ? S->cw2[S->cid[i]] : 0.0);
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
          if (y != x && S->ng && S->cid[i] == S->cid[j]) M[(size_t)x * m + y] += S->cw2[S->cid[i]];
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

// S_i = s_i R_i (row-major 3x3), s_i = sum |TV_j - TV_i| / (sum |V_j - V_i| + 1e-8)
__global__ void k_rot_frames(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ V, const double* __restrict__ TV, double* __restrict__ SR) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double lo = 0, lc = 0;
  const double v0 = V[3 * i], v1 = V[3 * i + 1], v2 = V[3 * i + 2];
  const double t0 = TV[3 * i], t1 = TV[3 * i + 1], t2 = TV[3 * i + 2];
  for (int e = rp[i]; e < rp[i + 1]; ++e) {
    const int j = ci[e];
    const double d1[3] = {V[3 * j] - v0, V[3 * j + 1] - v1, V[3 * j + 2] - v2};
    const double d2[3] = {TV[3 * j] - t0, TV[3 * j + 1] - t1, TV[3 * j + 2] - t2};
    lo += sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    lc += sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) C[3 * r + c] += d2[r] * d1[c];
  }
  const double sc = lc / (lo + 1e-8);
  double R[9];
  polar3(C, R);
  for (int k = 0; k < 9; ++k) SR[9 * (size_t)i + k] = sc * R[k];
}

// b_i = TV_i + sum_j (w_ij / 2)(S_i + S_j)(V_i - V_j) over the neighbours
// of vertex i (w_ij = 0 for a coincident neighbour)
__global__ void k_rot_rhs(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ w,
                          const double* __restrict__ V, const double* __restrict__ TV, const double* __restrict__ SR,
                          double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc[3] = {TV[3 * i], TV[3 * i + 1], TV[3 * i + 2]};
  const double* Si = SR + 9 * (size_t)i;
  for (int e = rp[i]; e < rp[i + 1]; ++e) {
    const double we = w[e];
    if (we == 0.0) continue;
    const int j = ci[e];
    const double h = 0.5 * we;
    const double* Sj = SR + 9 * (size_t)j;
    const double d[3] = {V[3 * i] - V[3 * j], V[3 * i + 1] - V[3 * j + 1], V[3 * i + 2] - V[3 * j + 2]};
    for (int r = 0; r < 3; ++r)
      acc[r] += h * ((Si[3 * r] + Sj[3 * r]) * d[0] + (Si[3 * r + 1] + Sj[3 * r + 1]) * d[1] +
                     (Si[3 * r + 2] + Sj[3 * r + 2]) * d[2]);
  }
  for (int r = 0; r < 3; ++r) b[3 * i + r] = acc[r];
}

// B_c = sum of b_i over the members of class c (in index order); the
// initial guess is TV of its first member
__global__ void k_rot_gather(int nc, const int* __restrict__ mp, const int* __restrict__ mm,
                             const double* __restrict__ b, const double* __restrict__ TV, double* __restrict__ B,
                             double* __restrict__ X) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double acc[3] = {0, 0, 0};
  for (int k = mp[c]; k < mp[c + 1]; ++k) {
    const int i = mm[k];
    for (int r = 0; r < 3; ++r) acc[r] += b[3 * i + r];
  }
  const int f = mm[mp[c]];
  for (int r = 0; r < 3; ++r) {
    B[3 * c + r] = acc[r];
    X[3 * c + r] = TV[3 * f + r];
  }
}

__global__ void k_rot_scatter(int n, const int* __restrict__ cls, const double* __restrict__ X,
                              double* __restrict__ V) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cls[i];
  for (int r = 0; r < 3; ++r) V[3 * i + r] = X[3 * c + r];
}

// Uploads the system, lets `prep` fill device B (or uploads S.B when prep is
// empty) and runs the CG kernel on X (initial guess in, solution out).
template <class Prep>
int run_cg(const System& S, double* X, int max_iter, double rtol, int* iterations, hipStream_t s, const char* who,
           Prep prep) {
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t nn = (size_t)S.n, ng = (size_t)S.ng, nnz = S.ci.size(), nb = S.bmem.size(), ni = S.binv.size();
  const size_t bytes = al((nn + 1) * 4) + al(nnz * 4) + al(nnz * 8) + al(nn * 4) + al((ng + 1) * 4) + al(nn * 4) +
                       al(ng * 8) + 3 * al(nn * 4) + al(nb * 4) + al(ni * 8) + 5 * al(nn * 24) + al(ng * 24) +
                       al(32) + al(nn * 8) + al(2 * kMaxTeam * 8 * 8);
  char* ws = nullptr;
  MO_HIP(hipMallocAsync((void**)&ws, bytes, s));
  char* w = ws;
  auto take = [&](size_t b) {
    char* p = w;
    w += al(b);
    return p;
  };
  CGArgs a;
  a.n = S.n;
  a.ng = S.ng;
  a.c0 = S.c0;
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
  a.info = (int*)take(32);   // info[0..1]; team counter [4], abort flag [5]
  double* d_dg = (double*)take(S.dg.size() * 8);
  double* d_part = (double*)take(2 * kMaxTeam * 8 * 8);   // team partials [2][kMaxTeam][8]
  a.dg = S.dg.empty() ? nullptr : d_dg;
  a.rp = d_rp; a.ci = d_ci; a.w = d_w; a.cid = d_cid; a.cp = d_cp; a.cm = d_cm; a.cw2 = d_cw2;
  a.pstart = d_ps; a.psize = d_pz; a.prow = d_pr; a.bmem = d_bm; a.binv = d_bi; a.B = d_B;
  a.X = X;
  a.max_iter = max_iter;
  a.rtol = rtol;
  int rc = MO_OK;
  do {
    struct Up { void* d; const void* h; size_t b; } ups[] = {
        {d_rp, S.rp.data(), (nn + 1) * 4}, {d_ci, S.ci.data(), nnz * 4}, {d_w, S.w.data(), nnz * 8},
        {d_cid, S.cid.data(), S.cid.size() * 4}, {d_cp, S.cp.data(), S.cp.size() * 4},
        {d_cm, S.cm.data(), S.cm.size() * 4}, {d_cw2, S.cw2.data(), ng * 8}, {d_ps, S.pstart.data(), nn * 4},
        {d_pz, S.psize.data(), nn * 4}, {d_pr, S.prow.data(), nn * 4}, {d_bm, S.bmem.data(), nb * 4},
        {d_bi, S.binv.data(), ni * 8}, {d_B, S.B.data(), S.B.size() * 8}, {d_dg, S.dg.data(), S.dg.size() * 8}};
    bool ok = true;
    for (const Up& u : ups)
      if (u.b && hipMemcpyAsync(u.d, u.h, u.b, hipMemcpyHostToDevice, s) != hipSuccess) ok = false;
    if (!ok) {
      rc = fail(MO_ERR_HIP, "%s: upload failed", who);
      break;
    }
    if (!prep(a, d_B)) {
      rc = fail(MO_ERR_HIP, "%s: set-up kernels failed", who);
      break;
    }
    // Systems of at least kTeamRows rows run on a cooperative grid of
    // ceil(n / kCGThreads) workgroups (at most lm_team_max(), at most what is
    // co-resident), one such launch at a time in the process.
    int team = 1;
    if (S.n >= kTeamRows) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_cg3_team, kCGThreads, 0) == hipSuccess)
        team = std::min({(S.n + kCGThreads - 1) / kCGThreads, lm_team_max(), per * cus});
      team = std::max(team, 1);
    }
    std::unique_lock<std::mutex> team_lock(team_launch_mutex(), std::defer_lock);
    if (team > 1) {
      int dev = 0, khz = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) {
        rc = fail(MO_ERR_HIP, "%s: device query failed", who);
        break;
      }
      long long limit = (long long)std::max(khz, 1) * 1000LL * 20;   // a 20 s bound on any one wait
      if (const char* wv = getenv("MO_LM_TEAM_WAIT_US")) limit = atoll(wv) * (long long)std::max(khz, 1) / 1000LL;
      unsigned* cnt = (unsigned*)a.info + 4;                  // counter + abort flag after info[0..1]
      int* abort_flag = (int*)a.info + 5;
      double* part = d_part;
      hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, a.info, 8);
      if (hipGetLastError() != hipSuccess) {
        rc = fail(MO_ERR_HIP, "%s: clear failed", who);
        break;
      }
      void* args[] = {(void*)&a, (void*)&cnt, (void*)&part, (void*)&abort_flag, (void*)&limit};
      team_lock.lock();
      if (hipLaunchCooperativeKernel((const void*)k_cg3_team, dim3(team), dim3(kCGThreads), args, 0, s) !=
          hipSuccess) {
        (void)hipGetLastError();   // refused: nothing was launched; solve on one workgroup
        team = 1;
      }
    }
    if (team == 1) {
      hipLaunchKernelGGL(k_cg3, dim3(1), dim3(kCGThreads), 0, s, a);
      if (hipGetLastError() != hipSuccess) {
        rc = fail(MO_ERR_HIP, "%s: launch failed", who);
        break;
      }
    }
    int info[6] = {0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(info, a.info, 24, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(MO_ERR_HIP, "%s: solve failed", who);
      break;
    }
    if (team > 1 && info[5]) {
      rc = fail(MO_ERR_HIP, "%s: a team barrier wait exceeded its bound (result invalid)", who);
      break;
    }
    if (iterations) *iterations = info[0];
    if (!info[1]) rc = fail(MO_ERR_NOCONV, "%s: CG did not reach rtol %g in %d iterations", who, rtol, info[0]);
  } while (0);
  (void)hipFreeAsync(ws, s);
  return rc;
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
  // the rest positions are the initial guess
  return run_cg(S, V, max_iter, rtol, iterations, s, "linear_estimation",
                [](const CGArgs&, double*) { return true; });
}

int linear_estimation_rot(double* V, int64_t nv, const int32_t* F, int64_t nf, const double* TV, double rigidity,
                          int max_iter, double rtol, int* iterations, hipStream_t s) {
  MO_CHECK(nv > 0 && nv < (1LL << 30), MO_ERR_ARG, "linear_estimation_rot: bad vertex count %lld", (long long)nv);
  MO_CHECK(nf >= 0 && nf < (1LL << 29), MO_ERR_ARG, "linear_estimation_rot: bad face count %lld", (long long)nf);
  const int n = (int)nv;
  std::vector<double> hV(nv * 3);
  std::vector<int32_t> hF(nf * 3);
  MO_HIP(hipMemcpyAsync(hV.data(), V, hV.size() * 8, hipMemcpyDeviceToHost, s));
  if (nf) MO_HIP(hipMemcpyAsync(hF.data(), F, hF.size() * 4, hipMemcpyDeviceToHost, s));
  MO_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < nf * 3; ++i)
    MO_CHECK(hF[i] >= 0 && hF[i] < nv, MO_ERR_ARG, "linear_estimation_rot: face index out of range");
  // Per face half-edge (a,b) and direction: r = 2e-2 / |V_b - V_a| * rigidity
  // (linear.cc:185-202); an undirected pair joined by m half-edges has the
  // Laplacian weight 2 m r.  Coincident endpoints (|e| = 0, which Subdivide
  // produces where a lattice point falls on a split edge) would make r
  // infinite and the reference's right-hand side NaN; such vertices are
  // solved as one unknown, the limit of the system as |e| -> 0.
  std::vector<int> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  std::unordered_map<uint64_t, double> vw;  // vertex pairs -> 2 m r (0 when coincident)
  vw.reserve((size_t)nf * 6);
  for (int64_t f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      const int a = hF[3 * f + j], b = hF[3 * f + (j + 1) % 3];
      if (a == b) continue;
      const double d0 = hV[3 * b] - hV[3 * a], d1 = hV[3 * b + 1] - hV[3 * a + 1], d2 = hV[3 * b + 2] - hV[3 * a + 2];
      const double len = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      const uint64_t k = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
      double& w = vw[k];
      if (len > 0.0) {
        w += 2.0 * (2e-2 / len * rigidity);
      } else {
        const int ra = find(a), rb = find(b);
        if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
      }
    }
  // classes numbered by their lowest member; members in index order
  std::vector<int> cls(n), mp, mm(n);
  int nc = 0;
  {
    std::vector<int> id(n, -1);
    for (int i = 0; i < n; ++i) {
      const int r = find(i);
      if (id[r] < 0) id[r] = nc++;
      cls[i] = id[r];
    }
    mp.assign(nc + 1, 0);
    for (int i = 0; i < n; ++i) mp[cls[i] + 1]++;
    for (int c = 0; c < nc; ++c) mp[c + 1] += mp[c];
    std::vector<int> fill(mp.begin(), mp.end() - 1);
    for (int i = 0; i < n; ++i) mm[fill[cls[i]]++] = i;
  }
  // vertex neighbour rows (all face-edge neighbours, sorted: the frames'
  // std::set order) with their weights
  std::vector<std::pair<uint64_t, double>> ed(vw.begin(), vw.end());
  std::sort(ed.begin(), ed.end());
  std::vector<int> orp(n + 1, 0), oci;
  std::vector<double> ow;
  for (const auto& e : ed) {
    orp[(e.first >> 32) + 1]++;
    orp[(uint32_t)e.first + 1]++;
  }
  for (int i = 0; i < n; ++i) orp[i + 1] += orp[i];
  oci.resize(orp[n]);
  ow.resize(orp[n]);
  {
    std::vector<int> fill(orp.begin(), orp.end() - 1);
    for (const auto& e : ed) {  // keys ascending: each row comes out sorted
      const int a = (int)(e.first >> 32), b = (int)(uint32_t)e.first;
      oci[fill[a]] = b;
      ow[fill[a]++] = e.second;
    }
    for (const auto& e : ed) {
      const int a = (int)(e.first >> 32), b = (int)(uint32_t)e.first;
      oci[fill[b]] = a;
      ow[fill[b]++] = e.second;
    }
    for (int i = 0; i < n; ++i) {
      std::vector<std::pair<int, double>> row;
      for (int e = orp[i]; e < orp[i + 1]; ++e) row.emplace_back(oci[e], ow[e]);
      std::sort(row.begin(), row.end());
      for (size_t k = 0; k < row.size(); ++k) {
        oci[orp[i] + k] = row[k].first;
        ow[orp[i] + k] = row[k].second;
      }
    }
  }
  // the class system: unit pull per member, Laplacian between classes
  System S;
  S.n = nc;
  S.ng = 0;
  S.c0 = 1.0;
  if (nc != n) {
    S.dg.resize(nc);
    for (int c = 0; c < nc; ++c) S.dg[c] = (double)(mp[c + 1] - mp[c]);
  }
  std::unordered_map<uint64_t, double> wsum;
  wsum.reserve(ed.size());
  for (const auto& e : ed) {
    if (e.second == 0.0) continue;
    const int a = cls[(int)(e.first >> 32)], b = cls[(int)(uint32_t)e.first];
    if (a == b) continue;
    wsum[((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b)] += e.second;
  }
  finish_system(wsum, &S);

  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t nn = (size_t)n, nnz = oci.size();
  const size_t bytes = al((nn + 1) * 4) + al(nnz * 4) + al(nnz * 8) + al(nn * 4) + al(((size_t)nc + 1) * 4) +
                       al(nn * 4) + al(nn * 24) + al(nn * 72) + al(nn * 24) + al((size_t)nc * 24);
  char* ws = nullptr;
  MO_HIP(hipMallocAsync((void**)&ws, bytes, s));
  char* w = ws;
  auto take = [&](size_t b) {
    char* p = w;
    w += al(b);
    return p;
  };
  int* d_orp = (int*)take((nn + 1) * 4);
  int* d_oci = (int*)take(nnz * 4);
  double* d_ow = (double*)take(nnz * 8);
  int* d_cls = (int*)take(nn * 4);
  int* d_mp = (int*)take(((size_t)nc + 1) * 4);
  int* d_mm = (int*)take(nn * 4);
  double* d_V0 = (double*)take(nn * 24);
  double* d_SR = (double*)take(nn * 72);
  double* d_b = (double*)take(nn * 24);
  double* d_X = (double*)take((size_t)nc * 24);
  int rc = MO_OK;
  do {
    struct Up { void* d; const void* h; size_t b; } ups[] = {
        {d_orp, orp.data(), (nn + 1) * 4}, {d_oci, oci.data(), nnz * 4}, {d_ow, ow.data(), nnz * 8},
        {d_cls, cls.data(), nn * 4}, {d_mp, mp.data(), ((size_t)nc + 1) * 4}, {d_mm, mm.data(), nn * 4}};
    bool ok = hipMemcpyAsync(d_V0, V, nn * 24, hipMemcpyDeviceToDevice, s) == hipSuccess;
    for (const Up& u : ups)
      if (u.b && hipMemcpyAsync(u.d, u.h, u.b, hipMemcpyHostToDevice, s) != hipSuccess) ok = false;
    if (!ok) {
      rc = fail(MO_ERR_HIP, "linear_estimation_rot: upload failed");
      break;
    }
    const unsigned bn = (unsigned)((n + 255) / 256), bc = (unsigned)((nc + 255) / 256);
    rc = run_cg(S, d_X, max_iter, rtol, iterations, s, "linear_estimation_rot", [&](const CGArgs&, double* dB) {
      hipLaunchKernelGGL(k_rot_frames, dim3(bn), dim3(256), 0, s, n, d_orp, d_oci, d_V0, TV, d_SR);
      hipLaunchKernelGGL(k_rot_rhs, dim3(bn), dim3(256), 0, s, n, d_orp, d_oci, d_ow, d_V0, TV, d_SR, d_b);
      hipLaunchKernelGGL(k_rot_gather, dim3(bc), dim3(256), 0, s, nc, d_mp, d_mm, d_b, TV, dB, d_X);
      return hipGetLastError() == hipSuccess;
    });
    if (rc != MO_OK) break;
    hipLaunchKernelGGL(k_rot_scatter, dim3(bn), dim3(256), 0, s, n, d_cls, d_X, V);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      rc = fail(MO_ERR_HIP, "linear_estimation_rot: scatter failed");
  } while (0);
  (void)hipFreeAsync(ws, s);
  return rc;
}

}  // namespace mo
