// MeshODE deformation hot path — CPU restatement used as the ORACLE.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (meshode_amd/, the C-ABI
// library, the CLIs) links, loads or calls this file.  Only tests/, the smoke
// check in __graft_entry__.py and the cpu_baseline leg of bench.py load the
// shared library built from it, and only as the checker or the timed CPU
// baseline — never as a fallback for the HIP path.
//
// Every function restates the reference's arithmetic from its source (cited
// file:line, relative to the reference checkout).  Build flags matter:
// -O2 -ffp-contract=off on x86-64 (SSE2 doubles, no FMA), so that every
// float/double operation below is one IEEE-rounded operation in the order
// written.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   normalize / apply_transform / (de)normalize_by_template
//       closed form: unit-box placement, round trip, hand-computed cases.
//   grid build (point-triangle distance, canonical shared features)
//       closed forms for face/edge/vertex regions of one triangle, sphere
//       distance |r - R| bounds on an icosphere, 1-Lipschitz invariant,
//       brute force over an independent numpy formula, lowest-index ties.
//   lookup f32 / f64 (+ Jet gradients)
//       affine fields reproduced exactly, the >0.2 truncation, the
//       out-of-box penalties and the index N-1 zero-penalty quirk,
//       truncation toward zero, central differences away from kinks.
//   rigid / graph edge losses
//       closed form, zero at rest, translation invariance, sum of
//       gradients = 0, central differences of the summed loss.
//   LM (Levenberg-Marquardt restating Ceres' trust-region defaults)
//       limits: lambda=0 decouples vertices; single-vertex problems reach
//       the trilinear minimum; exact solve residual; monotone cost.
//       The exact Ceres iterate sequence is NOT pinned (Ceres is absent):
//       "parity unpinned" for iterate-level agreement with Ceres.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <vector>

#define OR_API extern "C" __attribute__((visibility("default")))

namespace {

// ---------------------------------------------------------------------------
// Float -> int conversion with x86 semantics (cvttss2si / cvttsd2si): values
// outside int32 range and NaN become INT_MIN.  The reference's
// `int px = p * N` is undefined there in C++; on its x86 build it yields
// INT_MIN, which then takes the "px < 0" penalty branch.
inline int trunc_i32(double f) {
  if (!(f > -2147483649.0 && f < 2147483648.0)) return INT_MIN;
  return (int)f;
}
inline int trunc_i32f(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
  return (int)f;
}

// ---------------------------------------------------------------------------
// Jet<T,3>: forward-mode dual number with the operation formulas of Ceres'
// jet.h (the reference's autodiff type; uniformgrid.cc:152-158 instantiates
// the lookups on Jet<double,3> and Jet<float,3>).
//   (f*g).a = f.a*g.a ; (f*g).v = f.a*g.v + f.v*g.a
//   (f+g) = (f.a+g.a, f.v+g.v) ; (f-g) = (f.a-g.a, f.v-g.v) ; -f = (-f.a,-f.v)
template <class T>
struct Jet {
  T a;
  T v[3];
  Jet() : a(0) { v[0] = v[1] = v[2] = 0; }
  explicit Jet(T s) : a(s) { v[0] = v[1] = v[2] = 0; }
  Jet(T s, int k) : a(s) { v[0] = v[1] = v[2] = 0; v[k] = 1; }
};
template <class T> inline Jet<T> operator*(const Jet<T>& f, const Jet<T>& g) {
  Jet<T> r; r.a = f.a * g.a;
  for (int k = 0; k < 3; ++k) { T x = f.a * g.v[k]; T y = f.v[k] * g.a; r.v[k] = x + y; }
  return r;
}
template <class T> inline Jet<T> operator+(const Jet<T>& f, const Jet<T>& g) {
  Jet<T> r; r.a = f.a + g.a;
  for (int k = 0; k < 3; ++k) r.v[k] = f.v[k] + g.v[k];
  return r;
}
template <class T> inline Jet<T> operator-(const Jet<T>& f, const Jet<T>& g) {
  Jet<T> r; r.a = f.a - g.a;
  for (int k = 0; k < 3; ++k) r.v[k] = f.v[k] - g.v[k];
  return r;
}
template <class T> inline Jet<T> operator-(const Jet<T>& f) {
  Jet<T> r; r.a = -f.a;
  for (int k = 0; k < 3; ++k) r.v[k] = -f.v[k];
  return r;
}


// scalar part of a number or Jet (Ceres' Jet comparisons compare .a only).
inline float scalar_of(float x) { return x; }
inline double scalar_of(double x) { return x; }
template <class T> inline T scalar_of(const Jet<T>& j) { return j.a; }

// ---------------------------------------------------------------------------
// UniformGrid lookups.  grid[(z*N + y)*N + x] holds voxel_distance_[z][y][x]
// (uniformgrid.h:33; filled by mesh.cc:143-150 with point (x/N, y/N, z/N)).

// UniformGrid::DistanceFloat<T> (uniformgrid.cc:85-150): FP32 arithmetic,
// corner values are the FP64 voxels cast to float at read (:121-143).
template <class J>
J distance_float_t(const double* grid, int N, const J* p, float pa0, float pa1, float pa2) {
  typedef J T;
  const int px = trunc_i32f(pa0 * (float)N);          // :87 (float * int)
  const int py = trunc_i32f(pa1 * (float)N);
  const int pz = trunc_i32f(pa2 * (float)N);
  if (px < 0 || py < 0 || pz < 0 || px >= N - 1 || py >= N - 1 || pz >= N - 1) {   // :90-93
    T l = T(0.0f);
    const T fN = T((float)N);
    const T edge = T((float)(N - 1 - 1e-3));             // (T)(N - 1 - 1e-3)
    if (px < 0) l = l + (-p[0]) * fN;                    // :96-100
    else if (px >= N) l = l + (p[0] * fN - edge);
    if (py < 0) l = l + (-p[1]) * fN;                    // :102-106
    else if (py >= N) l = l + (p[1] * fN - edge);
    if (pz < 0) l = l + (-p[2]) * fN;                    // :108-112
    else if (pz >= N) l = l + (p[2] * fN - edge);
    return l;                                            // :114 (no truncation)
  }
  const T fN = T((float)N);
  const T one = T(1.0f);
  T wx = p[0] * fN - T((float)px);                       // :117-119
  T wy = p[1] * fN - T((float)py);
  T wz = p[2] * fN - T((float)pz);
  auto V = [&](int z, int y, int x) { return T((float)grid[((size_t)z * N + y) * N + x]); };
  T w0 = (one - wx) * (one - wy) * (one - wz) * V(pz, py, px);           // :121-143
  T w1 = wx * (one - wy) * (one - wz) * V(pz, py, px + 1);
  T w2 = (one - wx) * wy * (one - wz) * V(pz, py + 1, px);
  T w3 = wx * wy * (one - wz) * V(pz, py + 1, px + 1);
  T w4 = (one - wx) * (one - wy) * wz * V(pz + 1, py, px);
  T w5 = wx * (one - wy) * wz * V(pz + 1, py, px + 1);
  T w6 = (one - wx) * wy * wz * V(pz + 1, py + 1, px);
  T w7 = wx * wy * wz * V(pz + 1, py + 1, px + 1);
  T res = w0 + w1 + w2 + w3 + w4 + w5 + w6 + w7;         // :145
  if (scalar_of(res) > 0.2f) return T(0.0f);             // :147-149
  return res;
}

// UniformGrid::distance<T> (uniformgrid.cc:18-83): FP64 arithmetic.
template <class J>
J distance_double_t(const double* grid, int N, const J* p, double pa0, double pa1, double pa2) {
  typedef J T;
  const int px = trunc_i32(pa0 * N);                     // :20-22 (double * int)
  const int py = trunc_i32(pa1 * N);
  const int pz = trunc_i32(pa2 * N);
  if (px < 0 || py < 0 || pz < 0 || px >= N - 1 || py >= N - 1 || pz >= N - 1) {   // :23-26
    T l = T(0.0);
    const T fN = T((double)N);
    const T edge = T((double)(N - 1 - 1e-3));
    if (px < 0) l = l + (-p[0]) * fN;                    // :29-33
    else if (px >= N) l = l + (p[0] * fN - edge);
    if (py < 0) l = l + (-p[1]) * fN;
    else if (py >= N) l = l + (p[1] * fN - edge);
    if (pz < 0) l = l + (-p[2]) * fN;
    else if (pz >= N) l = l + (p[2] * fN - edge);
    return l;
  }
  const T fN = T((double)N);
  const T one = T(1.0);
  T wx = p[0] * fN - T((double)px);                      // :50-52
  T wy = p[1] * fN - T((double)py);
  T wz = p[2] * fN - T((double)pz);
  auto V = [&](int z, int y, int x) { return T(grid[((size_t)z * N + y) * N + x]); };
  T w0 = (one - wx) * (one - wy) * (one - wz) * V(pz, py, px);           // :54-76
  T w1 = wx * (one - wy) * (one - wz) * V(pz, py, px + 1);
  T w2 = (one - wx) * wy * (one - wz) * V(pz, py + 1, px);
  T w3 = wx * wy * (one - wz) * V(pz, py + 1, px + 1);
  T w4 = (one - wx) * (one - wy) * wz * V(pz + 1, py, px);
  T w5 = wx * (one - wy) * wz * V(pz + 1, py, px + 1);
  T w6 = (one - wx) * wy * wz * V(pz + 1, py + 1, px);
  T w7 = wx * wy * wz * V(pz + 1, py + 1, px + 1);
  T res = w0 + w1 + w2 + w3 + w4 + w5 + w6 + w7;         // :78
  if (scalar_of(res) > 0.2) return T(0.0);               // :80-82
  return res;
}


// ---------------------------------------------------------------------------
// Point-triangle squared distance, canonical form.
//
// The reference calls igl::point_mesh_squared_distance (mesh.cc:140): the
// unsigned squared Euclidean distance from each grid point to the nearest
// triangle (its index I and closest point C are discarded, mesh.cc:137-150).
// This restatement fixes the arithmetic so that results are reproducible to
// the bit and ties between triangles sharing an edge or a vertex are exact:
//   * face region: when the projection of q falls inside the triangle
//     (barycentric numerators bb, cc >= 0 and bb + cc <= denom), the plane
//     distance (n.(q-a))^2 * inv_nn;
//   * otherwise the minimum over the three edges, each evaluated with its
//     endpoints ordered by vertex index (lower index first), so every
//     triangle sharing that edge or vertex computes the identical value.
// Nearest index: lowest triangle index among equal squared distances.

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct FaceData {
  V3 p[3];        // positions in face order a, b, c
  int idx[3];     // vertex indices
  V3 n;           // (b-a) x (c-a)
  double inv_nn;  // 1 / (n.n)
  double d00, d01, d11, denom;
  double ll[3], inv_ll[3];   // canonical edge j: between p[j] and p[(j+1)%3]
};

inline FaceData make_face(const double* V, const int* F, long f) {
  FaceData d;
  for (int j = 0; j < 3; ++j) {
    int vi = F[f * 3 + j];
    d.idx[j] = vi;
    d.p[j] = V3{V[(size_t)vi * 3 + 0], V[(size_t)vi * 3 + 1], V[(size_t)vi * 3 + 2]};
  }
  V3 e0 = sub(d.p[1], d.p[0]);
  V3 e1 = sub(d.p[2], d.p[0]);
  d.n = cross(e0, e1);
  double nn = dot(d.n, d.n);
  d.inv_nn = nn > 0 ? 1.0 / nn : 0.0;
  d.d00 = dot(e0, e0);
  d.d01 = dot(e0, e1);
  d.d11 = dot(e1, e1);
  d.denom = d.d00 * d.d11 - d.d01 * d.d01;
  for (int j = 0; j < 3; ++j) {
    int i0 = j, i1 = (j + 1) % 3;
    if (d.idx[i1] < d.idx[i0]) std::swap(i0, i1);
    V3 dv = sub(d.p[i1], d.p[i0]);
    d.ll[j] = dot(dv, dv);
    d.inv_ll[j] = d.ll[j] > 0 ? 1.0 / d.ll[j] : 0.0;
  }
  return d;
}

inline double edge_d2(const FaceData& f, int j, V3 q) {
  int i0 = j, i1 = (j + 1) % 3;
  if (f.idx[i1] < f.idx[i0]) std::swap(i0, i1);
  V3 u = f.p[i0], w = f.p[i1];
  V3 dv = sub(w, u);
  V3 qu = sub(q, u);
  double tn = dot(qu, dv);
  if (tn <= 0.0 || f.ll[j] == 0.0) return dot(qu, qu);
  if (tn >= f.ll[j]) { V3 qw = sub(q, w); return dot(qw, qw); }
  double t = tn * f.inv_ll[j];
  V3 c = V3{u.x + t * dv.x, u.y + t * dv.y, u.z + t * dv.z};
  V3 qc = sub(q, c);
  return dot(qc, qc);
}

inline double point_face_d2(const FaceData& f, V3 q) {
  V3 v = sub(q, f.p[0]);
  if (f.denom > 0.0 && f.inv_nn > 0.0) {
    V3 e0 = sub(f.p[1], f.p[0]);
    V3 e1 = sub(f.p[2], f.p[0]);
    double d20 = dot(v, e0);
    double d21 = dot(v, e1);
    double bb = f.d11 * d20 - f.d01 * d21;
    double cc = f.d00 * d21 - f.d01 * d20;
    if (bb >= 0.0 && cc >= 0.0 && bb + cc <= f.denom) {
      double dp = dot(f.n, v);
      return (dp * dp) * f.inv_nn;
    }
  }
  double d0 = edge_d2(f, 0, q);
  double d1 = edge_d2(f, 1, q);
  double d2 = edge_d2(f, 2, q);
  return std::min(d0, std::min(d1, d2));
}

// ---------------------------------------------------------------------------
// Sparse symmetric positive definite solver for the LM normal equations:
// reverse Cuthill-McKee ordering at vertex level + envelope (skyline)
// Cholesky.  Stands in for Ceres' SPARSE_NORMAL_CHOLESKY (deformer.cc:57),
// which solves the same system exactly.
struct Skyline {
  int n = 0;
  std::vector<int> first;      // first column of row i (permuted dof order)
  std::vector<size_t> start;   // offset of row i in vals (entry for column first[i])
  std::vector<double> vals;
  double& at(int i, int j) { return vals[start[i] + (size_t)(j - first[i])]; }
  double get(int i, int j) const {
    if (j < first[i] || j > i) return 0.0;
    return vals[start[i] + (size_t)(j - first[i])];
  }
  bool factor() {
    for (int i = 0; i < n; ++i) {
      for (int j = first[i]; j <= i; ++j) {
        double s = vals[start[i] + (size_t)(j - first[i])];
        int k0 = std::max(first[i], first[j]);
        const double* li = &vals[start[i] + (size_t)(k0 - first[i])];
        const double* lj = &vals[start[j] + (size_t)(k0 - first[j])];
        for (int k = 0; k < j - k0; ++k) s -= li[k] * lj[k];
        if (j < i) {
          vals[start[i] + (size_t)(j - first[i])] = s / vals[start[j] + (size_t)(j - first[j])];
        } else {
          if (!(s > 0.0) || !std::isfinite(s)) return false;
          vals[start[i] + (size_t)(i - first[i])] = std::sqrt(s);
        }
      }
    }
    return true;
  }
  void solve(std::vector<double>& b) const {   // in place: L L^T x = b
    for (int i = 0; i < n; ++i) {
      double s = b[i];
      for (int k = first[i]; k < i; ++k) s -= vals[start[i] + (size_t)(k - first[i])] * b[k];
      b[i] = s / vals[start[i] + (size_t)(i - first[i])];
    }
    for (int i = n - 1; i >= 0; --i) {
      double xi = b[i] / vals[start[i] + (size_t)(i - first[i])];
      b[i] = xi;
      for (int k = first[i]; k < i; ++k) b[k] -= vals[start[i] + (size_t)(k - first[i])] * xi;
    }
  }
};

// Reverse Cuthill-McKee over a vertex adjacency (CSR).  Returns perm[v] = new
// position of vertex v.
std::vector<int> rcm_order(int nv, const std::vector<int>& rp, const std::vector<int>& ci) {
  std::vector<int> order; order.reserve(nv);
  std::vector<char> seen(nv, 0);
  std::vector<int> deg(nv);
  for (int v = 0; v < nv; ++v) deg[v] = rp[v + 1] - rp[v];
  std::vector<int> byDeg(nv);
  for (int v = 0; v < nv; ++v) byDeg[v] = v;
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  for (int s : byDeg) {
    if (seen[s]) continue;
    // find a pseudo-peripheral start by two BFS sweeps from s
    int start = s;
    for (int sweep = 0; sweep < 2; ++sweep) {
      std::vector<int> level(nv, -1);
      std::deque<int> q; q.push_back(start); level[start] = 0;
      int last = start;
      while (!q.empty()) {
        int u = q.front(); q.pop_front(); last = u;
        for (int k = rp[u]; k < rp[u + 1]; ++k) { int w = ci[k]; if (level[w] < 0) { level[w] = level[u] + 1; q.push_back(w); } }
      }
      start = last;
    }
    std::deque<int> q; q.push_back(start); seen[start] = 1;
    while (!q.empty()) {
      int u = q.front(); q.pop_front(); order.push_back(u);
      std::vector<int> nb;
      for (int k = rp[u]; k < rp[u + 1]; ++k) { int w = ci[k]; if (!seen[w]) { seen[w] = 1; nb.push_back(w); } }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
      for (int w : nb) q.push_back(w);
    }
  }
  std::reverse(order.begin(), order.end());
  std::vector<int> perm(nv);
  for (int i = 0; i < nv; ++i) perm[order[i]] = i;
  return perm;
}

}  // namespace

// ===========================================================================
// Exported API (ctypes).  Arrays are row-major, contiguous.

OR_API int mo_oracle_version() { return 1; }

// Mesh::Normalize (mesh.cc:66-96): scale = 1.1 * max axis extent,
// pos = min - 0.05 * scale, v <- (v - pos) / scale.
OR_API void mo_oracle_normalize(double* V, long n, double* scale_out, double* pos_out) {
  double mn[3], mx[3];
  for (int j = 0; j < 3; ++j) {
    mn[j] = 1e30; mx[j] = -1e30;
    for (long i = 0; i < n; ++i) {
      if (V[i * 3 + j] < mn[j]) mn[j] = V[i * 3 + j];
      if (V[i * 3 + j] > mx[j]) mx[j] = V[i * 3 + j];
    }
  }
  double scale = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2])) * 1.1;
  double pos[3];
  for (int j = 0; j < 3; ++j) pos[j] = mn[j] - 0.05 * scale;
  for (long i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) V[i * 3 + j] = (V[i * 3 + j] - pos[j]) / scale;
  *scale_out = scale;
  for (int j = 0; j < 3; ++j) pos_out[j] = pos[j];
}

// Mesh::ApplyTransform (mesh.cc:98-105): v <- (v - pos) / scale.
OR_API void mo_oracle_apply_transform(double* V, long n, double scale, const double* pos) {
  for (long i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) V[i * 3 + j] = (V[i * 3 + j] - pos[j]) / scale;
}

// NormalizeByTemplate (normalize.cc:5-24) on float32 storage: FP64 math,
// result rounded to float.
OR_API void mo_oracle_normalize_by_template_f32(float* V, long n, double scale, const double* trans) {
  for (long i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) V[i * 3 + j] = (float)(((double)V[i * 3 + j] - trans[j]) / scale);
}

// DenormalizeByTemplate (normalize.cc:26-45).
OR_API void mo_oracle_denormalize_by_template_f32(float* V, long n, double scale, const double* trans) {
  for (long i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) V[i * 3 + j] = (float)((double)V[i * 3 + j] * scale + trans[j]);
}

// Mesh::ReflectionSymmetrize (mesh.cc:234-246): append x-mirrored copies.
// V must have room for 2n vertices and F for 2m faces.
OR_API void mo_oracle_reflection_symmetrize(double* V, long n, int* F, long m) {
  for (long i = 0; i < n; ++i) {
    V[(n + i) * 3 + 0] = -V[i * 3 + 0];
    V[(n + i) * 3 + 1] = V[i * 3 + 1];
    V[(n + i) * 3 + 2] = V[i * 3 + 2];
  }
  for (long i = 0; i < m; ++i)
    for (int j = 0; j < 3; ++j) F[(m + i) * 3 + j] = F[i * 3 + j] + (int)n;
}

// Squared distance + nearest triangle for arbitrary query points.
OR_API void mo_oracle_point_mesh_d2(const double* Q, long nq, const double* V, const int* F, long nf,
                                    double* d2_out, int* idx_out) {
  std::vector<FaceData> faces(nf);
  for (long f = 0; f < nf; ++f) faces[f] = make_face(V, F, f);
  for (long i = 0; i < nq; ++i) {
    V3 q{Q[i * 3], Q[i * 3 + 1], Q[i * 3 + 2]};
    double best = INFINITY; int bi = -1;
    for (long f = 0; f < nf; ++f) {
      double d = point_face_d2(faces[f], q);
      if (d < best) { best = d; bi = (int)f; }
    }
    d2_out[i] = best; idx_out[i] = bi;
  }
}

// Mesh::ConstructDistanceField (mesh.cc:106-152): grid point (i,j,k) at
// (k/N, j/N, i/N) (:110-116); voxel[i][j][k] = sqrt(min squared distance)
// (:140-150).  Writes dist[(i*N+j)*N+k], its float32 rounding and the
// nearest triangle index (lowest index on ties).  Optional subset: when
// `cells` is non-null only those linear cell ids are evaluated (outputs are
// indexed by position in `cells`).
OR_API void mo_oracle_build_grid(const double* V, const int* F, long nf, int N,
                                 const long* cells, long ncells,
                                 double* dist, float* dist32, int* nearest) {
  std::vector<FaceData> faces(nf);
  for (long f = 0; f < nf; ++f) faces[f] = make_face(V, F, f);
  long total = cells ? ncells : (long)N * N * N;
  for (long t = 0; t < total; ++t) {
    long c = cells ? cells[t] : t;
    int k = (int)(c % N), j = (int)((c / N) % N), i = (int)(c / ((long)N * N));
    V3 q{(double)k / N, (double)j / N, (double)i / N};
    double best = INFINITY; int bi = -1;
    for (long f = 0; f < nf; ++f) {
      double d = point_face_d2(faces[f], q);
      if (d < best) { best = d; bi = (int)f; }
    }
    double s = std::sqrt(best);
    if (dist) dist[t] = s;
    if (dist32) dist32[t] = (float)s;
    if (nearest) nearest[t] = bi;
  }
}

// UniformGrid::DistanceFloat<float> and <Jet<float,3>>: value and Jet
// derivative for each point.
OR_API void mo_oracle_lookup_f32(const double* grid, int N, const float* P, long n, float* d_out, float* g_out) {
  for (long i = 0; i < n; ++i) {
    const float* p = P + i * 3;
    float d = distance_float_t<float>(grid, N, p, p[0], p[1], p[2]);
    if (d_out) d_out[i] = d;
    if (g_out) {
      Jet<float> pj[3] = {Jet<float>(p[0], 0), Jet<float>(p[1], 1), Jet<float>(p[2], 2)};
      Jet<float> r = distance_float_t<Jet<float>>(grid, N, pj, p[0], p[1], p[2]);
      for (int k = 0; k < 3; ++k) g_out[i * 3 + k] = r.v[k];
    }
  }
}

// UniformGrid::distance<double> and <Jet<double,3>>.
OR_API void mo_oracle_lookup_f64(const double* grid, int N, const double* P, long n, double* d_out, double* g_out) {
  for (long i = 0; i < n; ++i) {
    const double* p = P + i * 3;
    double d = distance_double_t<double>(grid, N, p, p[0], p[1], p[2]);
    if (d_out) d_out[i] = d;
    if (g_out) {
      Jet<double> pj[3] = {Jet<double>(p[0], 0), Jet<double>(p[1], 1), Jet<double>(p[2], 2)};
      Jet<double> r = distance_double_t<Jet<double>>(grid, N, pj, p[0], p[1], p[2]);
      for (int k = 0; k < 3; ++k) g_out[i * 3 + k] = r.v[k];
    }
  }
}

// DistanceFieldLoss_forward (distance_layer.cc:8-37): out[i] = d(v_i)^2 (FP32).
OR_API void mo_oracle_distance_loss_fwd(const double* grid, int N, const float* V, long n, float* out) {
  for (long i = 0; i < n; ++i) {
    const float* p = V + i * 3;
    float d = distance_float_t<float>(grid, N, p, p[0], p[1], p[2]);
    out[i] = d * d;   // dataL[i] *= dataL[i]
  }
}

// DistanceFieldLoss_backward (distance_layer.cc:39-80): Jet<float,3> seeded
// with the identity; vd = d*d (Jet product); out = vd.v * 0.5 (double
// multiply, stored to float).
OR_API void mo_oracle_distance_loss_bwd(const double* grid, int N, const float* V, long n, float* out) {
  for (long i = 0; i < n; ++i) {
    const float* p = V + i * 3;
    Jet<float> pj[3] = {Jet<float>(p[0], 0), Jet<float>(p[1], 1), Jet<float>(p[2], 2)};
    Jet<float> vd = distance_float_t<Jet<float>>(grid, N, pj, p[0], p[1], p[2]);
    vd = vd * vd;
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = (float)((double)vd.v[k] * 0.5);
  }
}

// StoreRigidityInformation (rigid_layer.cc:11-46): per half-edge (f,j),
// offset = V[F[f][(j+1)%3]] - V[F[f][j]] (FP32).
OR_API void mo_oracle_store_rigidity(const float* V, const int* F, long nf, float* off) {
  for (long i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) {
      int v0 = F[i * 3 + j], v1 = F[i * 3 + (j + 1) % 3];
      for (int k = 0; k < 3; ++k) off[(i * 3 + j) * 3 + k] = V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k];
    }
}

// RigidEdgeLoss_forward (rigid_layer.cc:48-89).
OR_API void mo_oracle_rigid_fwd(const float* V, const int* F, long nf, const float* off, float* out) {
  for (long i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) {
      int v0 = F[i * 3 + j], v1 = F[i * 3 + (j + 1) % 3];
      for (int k = 0; k < 3; ++k) {
        float l = V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[(i * 3 + j) * 3 + k];
        out[(i * 3 + j) * 3 + k] = l * l;
      }
    }
}

// RigidEdgeLoss_backward (rigid_layer.cc:91-133): sequential scatter in face
// order, starting from zero.
OR_API void mo_oracle_rigid_bwd(const float* V, long nv, const int* F, long nf, const float* off, float* out) {
  std::memset(out, 0, sizeof(float) * 3 * nv);
  for (long i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) {
      int v0 = F[i * 3 + j], v1 = F[i * 3 + (j + 1) % 3];
      for (int k = 0; k < 3; ++k) {
        out[(size_t)v0 * 3 + k] -= (V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[(i * 3 + j) * 3 + k]);
        out[(size_t)v1 * 3 + k] += (V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[(i * 3 + j) * 3 + k]);
      }
    }
}

// StoreGraphInformation (graph_layer.cc:11-45): offset = V[e1] - V[e0].
OR_API void mo_oracle_store_graph(const float* V, const int* E, long ne, float* off) {
  for (long i = 0; i < ne; ++i) {
    int v0 = E[i * 2], v1 = E[i * 2 + 1];
    for (int k = 0; k < 3; ++k) off[i * 3 + k] = V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k];
  }
}

// GraphEdgeLoss_forward (graph_layer.cc:47-91).
OR_API void mo_oracle_graph_fwd(const float* V, const int* E, long ne, const float* off, float* out) {
  for (long i = 0; i < ne; ++i) {
    int v0 = E[i * 2], v1 = E[i * 2 + 1];
    for (int k = 0; k < 3; ++k) {
      float l = V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[i * 3 + k];
      out[i * 3 + k] = l * l;
    }
  }
}

// GraphEdgeLoss_backward (graph_layer.cc:93-135).
OR_API void mo_oracle_graph_bwd(const float* V, long nv, const int* E, long ne, const float* off, float* out) {
  std::memset(out, 0, sizeof(float) * 3 * nv);
  for (long i = 0; i < ne; ++i) {
    int v0 = E[i * 2], v1 = E[i * 2 + 1];
    for (int k = 0; k < 3; ++k) {
      out[(size_t)v0 * 3 + k] -= V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[i * 3 + k];
      out[(size_t)v1 * 3 + k] += V[(size_t)v1 * 3 + k] - V[(size_t)v0 * 3 + k] - off[i * 3 + k];
    }
  }
}

// ---------------------------------------------------------------------------
// Levenberg-Marquardt solve of Deformer::Deform / Deformer::DeformGraph
// (deformer.cc:18-92, 370-442).
//
// Problem (deformer.cc:30-53, 394-403): one DistanceLoss block per vertex,
// residual [d(x_u), 0, 0] with d = UniformGrid::distance<double>
// (distanceloss.h:6-25), and one EdgeLoss block per directed edge (a, b),
// residual lambda * (x_a - x_b - v) with v = x0_a - x0_b taken at problem
// build time (edgeloss.h:8-33, deformer.cc:44).  Cost = 1/2 sum ||r||^2.
//
// Minimizer: Ceres' trust-region minimizer with the Levenberg-Marquardt
// strategy and the options left at their defaults by deformer.cc:55-59
// (max_num_iterations = 100, SPARSE_NORMAL_CHOLESKY = exact solve here),
// restated from Ceres' documented algorithm:
//   * Jacobi scaling s = 1 / (1 + sqrt(colnorm^2(J))) computed once at
//     iteration 0 and applied to J every iteration;
//   * LM diagonal D^2 = clamp(colnorm^2(J s), 1e-6, 1e32) / radius, solve
//     (Js^T Js + D^2) y = Js^T f, step = -y, delta = s * step;
//   * model_cost_change = -(J delta).(f + J delta / 2); a step with
//     model_cost_change <= 0 is invalid (5 consecutive -> failure);
//   * convergence: ||x - x_cand|| <= 1e-8 (||x|| + 1e-8) (parameter),
//     |cost - cand_cost| <= 1e-6 cost (function), ||J^T f||_inf <= 1e-10 after
//     a successful step (gradient), radius <= 1e-32;
//   * rho = (cost - cand_cost) / model_cost_change; accept if rho > 1e-3:
//     radius /= max(1/3, 1 - (2 rho - 1)^3) (capped at 1e16), decrease
//     factor reset to 2; otherwise radius /= factor, factor *= 2.
// Ceres itself is absent; the iterate sequence is not pinned against it.

struct OracleLMOptions {
  int max_num_iterations;
  int max_consecutive_invalid_steps;
  double function_tolerance;
  double gradient_tolerance;
  double parameter_tolerance;
  double min_relative_decrease;
  double initial_trust_region_radius;
  double max_trust_region_radius;
  double min_trust_region_radius;
  double min_lm_diagonal;
  double max_lm_diagonal;
};

struct OracleLMStats {
  double initial_cost;
  double final_cost;
  double vertex_cost;
  double rigidity_cost;
  int iterations;
  int successful_steps;
  int unsuccessful_steps;
  int termination;   // 0 = convergence, 1 = no convergence (iteration cap), 2 = failure
};

namespace {

struct LMProblem {
  const double* grid; int N;
  int nv; int ne;
  const int* E;
  std::vector<double> rest;   // 3 per edge
  std::vector<double> lam;    // per-edge residual weight (EdgeLoss lambda, or AdaptiveEdgeLoss's)
};

double lm_cost(const LMProblem& P, const double* x, double* vcost, double* ecost) {
  double cv = 0.0, ce = 0.0;
  for (int u = 0; u < P.nv; ++u) {
    double d = distance_double_t<double>(P.grid, P.N, x + 3 * u, x[3 * u], x[3 * u + 1], x[3 * u + 2]);
    cv += 0.5 * d * d;
  }
  for (int h = 0; h < P.ne; ++h) {
    int a = P.E[2 * h], b = P.E[2 * h + 1];
    if (a == b) continue;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      double r = ((x[3 * a + k] - x[3 * b + k]) - P.rest[3 * h + k]) * P.lam[h];
      s += r * r;
    }
    ce += 0.5 * s;
  }
  if (vcost) *vcost = cv;
  if (ecost) *ecost = ce;
  return cv + ce;
}

// residual values d_u and gradients g_u of the distance blocks, J^T f and
// squared column norms of J.
void lm_jacobian(const LMProblem& P, const double* x, std::vector<double>& d, std::vector<double>& g,
                 std::vector<double>& grad, std::vector<double>& cn) {
  std::fill(grad.begin(), grad.end(), 0.0);
  std::fill(cn.begin(), cn.end(), 0.0);
  for (int u = 0; u < P.nv; ++u) {
    const double* p = x + 3 * u;
    Jet<double> pj[3] = {Jet<double>(p[0], 0), Jet<double>(p[1], 1), Jet<double>(p[2], 2)};
    Jet<double> r = distance_double_t<Jet<double>>(P.grid, P.N, pj, p[0], p[1], p[2]);
    d[u] = r.a;
    for (int k = 0; k < 3; ++k) {
      g[3 * u + k] = r.v[k];
      grad[3 * u + k] += r.a * r.v[k];
      cn[3 * u + k] += r.v[k] * r.v[k];
    }
  }
  for (int h = 0; h < P.ne; ++h) {
    int a = P.E[2 * h], b = P.E[2 * h + 1];
    if (a == b) continue;
    const double l = P.lam[h];
    for (int k = 0; k < 3; ++k) {
      double r = ((x[3 * a + k] - x[3 * b + k]) - P.rest[3 * h + k]) * l;
      grad[3 * a + k] += l * r;
      grad[3 * b + k] -= l * r;
      cn[3 * a + k] += l * l;
      cn[3 * b + k] += l * l;
    }
  }
}

}  // namespace

OR_API void mo_oracle_lm_default_options(OracleLMOptions* o) {
  o->max_num_iterations = 100;
  o->max_consecutive_invalid_steps = 5;
  o->function_tolerance = 1e-6;
  o->gradient_tolerance = 1e-10;
  o->parameter_tolerance = 1e-8;
  o->min_relative_decrease = 1e-3;
  o->initial_trust_region_radius = 1e4;
  o->max_trust_region_radius = 1e16;
  o->min_trust_region_radius = 1e-32;
  o->min_lm_diagonal = 1e-6;
  o->max_lm_diagonal = 1e32;
}

// Deformer::Deform / DeformGraph / DeformSubdivision.  x: [nv,3] FP64
// positions in the target's normalized frame (start point, updated in
// place); x0: the positions the rest offsets v_h = x0_a - x0_b are taken
// from (x itself when null).  E: [ne,2] directed edges (a,b) in residual
// order.  lam_e (optional, [ne]): per-edge residual weight; AdaptiveEdgeLoss
// (edgeloss.h:35-61) uses lambda * 2e-2 / (|v_h| + 1e-8), EdgeLoss the
// constant lambda (when lam_e is null).  max_successful > 0 stops after that
// many successful steps, as TerminateWhenSuccessCallback does after its
// second one (callback.h:22-36).  trace (optional, 8 doubles per iteration,
// up to max_num_iterations rows): iteration, cost, candidate cost, model
// cost change, rho, radius after the step, accepted flag, step norm.
OR_API int mo_oracle_lm_solve_ex(const double* grid, int N, double* x, const double* x0, int nv, const int* E,
                                 int ne, const double* lam_e, double lambda, int max_successful,
                                 const OracleLMOptions* opt, OracleLMStats* st, double* trace) {
  if (!x0) x0 = x;
  LMProblem P;
  P.grid = grid; P.N = N; P.nv = nv; P.ne = ne; P.E = E;
  P.rest.resize((size_t)3 * ne);
  for (int h = 0; h < ne; ++h)
    for (int k = 0; k < 3; ++k) P.rest[3 * h + k] = x0[3 * E[2 * h] + k] - x0[3 * E[2 * h + 1] + k];
  P.lam.assign(ne, lambda);
  if (lam_e) P.lam.assign(lam_e, lam_e + ne);

  const int n = 3 * nv;
  // vertex adjacency with multiplicities
  // vertex adjacency with summed squared edge weights
  std::vector<std::vector<std::pair<int, double>>> adj(nv);
  std::vector<double> deg(nv, 0.0);
  for (int h = 0; h < ne; ++h) {
    int a = E[2 * h], b = E[2 * h + 1];
    if (a == b) continue;
    const double w = P.lam[h] * P.lam[h];
    deg[a] += w; deg[b] += w;
    adj[a].push_back({b, w});
    adj[b].push_back({a, w});
  }
  std::vector<int> rp(nv + 1, 0), ci;
  std::vector<double> cw;
  for (int u = 0; u < nv; ++u) {
    auto& L = adj[u];
    std::stable_sort(L.begin(), L.end(), [](const std::pair<int, double>& p, const std::pair<int, double>& q) {
      return p.first < q.first;
    });
    std::vector<std::pair<int, double>> M;
    for (auto& e : L) {
      if (!M.empty() && M.back().first == e.first) M.back().second += e.second;
      else M.push_back(e);
    }
    L.swap(M);
    for (auto& e : L) { ci.push_back(e.first); cw.push_back(e.second); }
    rp[u + 1] = (int)ci.size();
  }
  std::vector<int> perm = rcm_order(nv, rp, ci);
  Skyline S;
  S.n = n;
  S.first.resize(n);
  S.start.resize(n + 1);
  for (int u = 0; u < nv; ++u) {
    int pu = perm[u], mn = pu;
    for (int k = rp[u]; k < rp[u + 1]; ++k) mn = std::min(mn, perm[ci[k]]);
    for (int k = 0; k < 3; ++k) S.first[3 * pu + k] = 3 * mn;
  }
  size_t tot = 0;
  for (int i = 0; i < n; ++i) { S.start[i] = tot; tot += (size_t)(i - S.first[i] + 1); }
  S.start[n] = tot;
  S.vals.assign(tot, 0.0);

  std::vector<double> d(nv), g(n), grad(n), cn(n), scale(n), diag(n), rhs(n), delta(n), cand(n);
  double cost = lm_cost(P, x, nullptr, nullptr);
  st->initial_cost = cost;
  lm_jacobian(P, x, d, g, grad, cn);
  for (int c = 0; c < n; ++c) scale[c] = 1.0 / (1.0 + std::sqrt(cn[c]));

  double radius = opt->initial_trust_region_radius, decrease = 2.0;
  bool reuse = false, last_success = true;
  int iter = 0, invalid = 0, nsucc = 0, nfail = 0, term = 1;
  auto gmax = [&]() { double m = 0; for (int c = 0; c < n; ++c) m = std::max(m, std::fabs(grad[c])); return m; };
  while (true) {
    if (iter >= opt->max_num_iterations) { term = 1; break; }
    if (last_success && gmax() <= opt->gradient_tolerance) { term = 0; break; }
    if (radius <= opt->min_trust_region_radius) { term = 0; break; }
    ++iter;
    if (!reuse)
      for (int c = 0; c < n; ++c)
        diag[c] = std::min(std::max(scale[c] * scale[c] * cn[c], opt->min_lm_diagonal), opt->max_lm_diagonal);
    reuse = true;
    // assemble S H S + D^2 in permuted order
    std::fill(S.vals.begin(), S.vals.end(), 0.0);
    for (int u = 0; u < nv; ++u) {
      int pu = perm[u];
      for (int k = 0; k < 3; ++k) {
        int row = 3 * pu + k;
        for (int l = 0; l <= k; ++l) {
          double h = g[3 * u + k] * g[3 * u + l] + (k == l ? deg[u] : 0.0);
          double v = scale[3 * u + k] * scale[3 * u + l] * h;
          if (k == l) v += diag[3 * u + k] / radius;
          S.at(row, 3 * pu + l) = v;
        }
        for (int e = rp[u]; e < rp[u + 1]; ++e) {
          int w = ci[e], pw = perm[w];
          if (pw < pu) S.at(row, 3 * pw + k) = -scale[3 * u + k] * scale[3 * w + k] * cw[e];
        }
      }
    }
    bool ok = S.factor();
    double mcc = -1.0;
    if (ok) {
      for (int u = 0; u < nv; ++u)
        for (int k = 0; k < 3; ++k) rhs[3 * perm[u] + k] = scale[3 * u + k] * grad[3 * u + k];
      S.solve(rhs);
      for (int u = 0; u < nv; ++u)
        for (int k = 0; k < 3; ++k) delta[3 * u + k] = scale[3 * u + k] * (-rhs[3 * perm[u] + k]);
      for (int c = 0; c < n; ++c) if (!std::isfinite(delta[c])) ok = false;
    }
    if (ok) {
      // model cost change = -(J delta).(f + J delta / 2)
      double acc = 0.0;
      for (int u = 0; u < nv; ++u) {
        double m = g[3 * u] * delta[3 * u] + g[3 * u + 1] * delta[3 * u + 1] + g[3 * u + 2] * delta[3 * u + 2];
        acc += m * (d[u] + 0.5 * m);
      }
      for (int h = 0; h < ne; ++h) {
        int a = E[2 * h], b = E[2 * h + 1];
        if (a == b) continue;
        const double l = P.lam[h];
        for (int k = 0; k < 3; ++k) {
          double r = ((x[3 * a + k] - x[3 * b + k]) - P.rest[3 * h + k]) * l;
          double m = l * (delta[3 * a + k] - delta[3 * b + k]);
          acc += m * (r + 0.5 * m);
        }
      }
      mcc = -acc;
    }
    if (!ok || !(mcc > 0.0)) {
      ++invalid;
      if (invalid >= opt->max_consecutive_invalid_steps) { term = 2; break; }
      radius /= decrease; decrease *= 2.0;
      last_success = false; ++nfail;
      if (trace) { double* t = trace + 8 * (iter - 1); t[0] = iter; t[1] = cost; t[2] = NAN; t[3] = mcc; t[4] = NAN; t[5] = radius; t[6] = 0; t[7] = NAN; }
      continue;
    }
    invalid = 0;
    for (int c = 0; c < n; ++c) cand[c] = x[c] + delta[c];
    double ccost = lm_cost(P, cand.data(), nullptr, nullptr);
    double sn = 0.0, xn = 0.0;
    for (int c = 0; c < n; ++c) { double t = x[c] - cand[c]; sn += t * t; xn += x[c] * x[c]; }
    sn = std::sqrt(sn); xn = std::sqrt(xn);
    double rho = (cost - ccost) / mcc;
    if (trace) { double* t = trace + 8 * (iter - 1); t[0] = iter; t[1] = cost; t[2] = ccost; t[3] = mcc; t[4] = rho; t[5] = radius; t[6] = 0; t[7] = sn; }
    if (sn <= opt->parameter_tolerance * (xn + opt->parameter_tolerance)) { term = 0; break; }
    if (std::fabs(cost - ccost) <= opt->function_tolerance * cost) { term = 0; break; }
    if (rho > opt->min_relative_decrease) {
      std::copy(cand.begin(), cand.end(), x);
      cost = ccost;
      lm_jacobian(P, x, d, g, grad, cn);
      radius = radius / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3));
      radius = std::min(opt->max_trust_region_radius, radius);
      decrease = 2.0; reuse = false; last_success = true; ++nsucc;
      if (trace) { double* t = trace + 8 * (iter - 1); t[5] = radius; t[6] = 1; }
      if (max_successful > 0 && nsucc >= max_successful) { term = 0; break; }
    } else {
      radius /= decrease; decrease *= 2.0; reuse = true; last_success = false; ++nfail;
      if (trace) { double* t = trace + 8 * (iter - 1); t[5] = radius; }
    }
  }
  double vc, ec;
  st->final_cost = lm_cost(P, x, &vc, &ec);
  st->vertex_cost = vc; st->rigidity_cost = ec;
  st->iterations = iter; st->successful_steps = nsucc; st->unsuccessful_steps = nfail;
  st->termination = term;
  return 0;
}

OR_API int mo_oracle_lm_solve(const double* grid, int N, double* x, int nv, const int* E, int ne,
                              double lambda, const OracleLMOptions* opt, OracleLMStats* st, double* trace) {
  return mo_oracle_lm_solve_ex(grid, N, x, nullptr, nv, E, ne, nullptr, lambda, 0, opt, st, trace);
}

// Cost terms of Deformer::Deform's final Evaluate calls (deformer.cc:76-91)
// at x with rest offsets taken from x0.
OR_API double mo_oracle_lm_cost_ex(const double* grid, int N, const double* x, const double* x0, int nv,
                                   const int* E, int ne, const double* lam_e, double lambda, double* vcost,
                                   double* ecost) {
  LMProblem P;
  P.grid = grid; P.N = N; P.nv = nv; P.ne = ne; P.E = E;
  P.lam.assign(ne, lambda);
  if (lam_e) P.lam.assign(lam_e, lam_e + ne);
  P.rest.resize((size_t)3 * ne);
  for (int h = 0; h < ne; ++h)
    for (int k = 0; k < 3; ++k) P.rest[3 * h + k] = x0[3 * E[2 * h] + k] - x0[3 * E[2 * h + 1] + k];
  return lm_cost(P, x, vcost, ecost);
}

OR_API double mo_oracle_lm_cost(const double* grid, int N, const double* x, const double* x0, int nv,
                                const int* E, int ne, double lambda, double* vcost, double* ecost) {
  return mo_oracle_lm_cost_ex(grid, N, x, x0, nv, E, ne, nullptr, lambda, vcost, ecost);
}

// ---------------------------------------------------------------------------
// Deformer::DeformWithRot (deformer.cc:94-171): unknowns x = positions [nv,3]
// and angle-axis rotations w [nv,3] (zero at the start, deformer.cc:114);
// residual blocks DistanceLoss(p_i) and, per face half-edge (a, b) =
// (F[f][j], F[f][j+1]) with a != b, EdgeLossWithRot (edgeloss.h:64-90):
//   r[0:3] = lambda (AngleAxisRotatePoint(w_a, p_a - p_b) - v),  v = x0_a - x0_b
//   r[3:6] = w_a - w_b.
// AngleAxisRotatePoint restates Ceres' published rotation.h (Ceres itself is
// absent): Rodrigues' formula when |w|^2 > DBL_EPSILON, else p + w x p.  The
// Jacobians are taken by forward-mode differentiation (JetK) of the same
// expressions, as Ceres' AutoDiffCostFunction does.  The LM loop and the
// exact sparse solve are those of mo_oracle_lm_solve_ex with 6 unknowns per
// vertex (p then w).
namespace {

template <int K>
struct JetK {
  double a;
  double v[K];
  JetK() : a(0) { for (int k = 0; k < K; ++k) v[k] = 0; }
  explicit JetK(double s) : a(s) { for (int k = 0; k < K; ++k) v[k] = 0; }
  JetK(double s, int i) : a(s) { for (int k = 0; k < K; ++k) v[k] = 0; v[i] = 1; }
};
template <int K> inline JetK<K> operator+(const JetK<K>& f, const JetK<K>& g) {
  JetK<K> r;
  r.a = f.a + g.a;
  for (int k = 0; k < K; ++k) r.v[k] = f.v[k] + g.v[k];
  return r;
}
template <int K> inline JetK<K> operator-(const JetK<K>& f, const JetK<K>& g) {
  JetK<K> r; r.a = f.a - g.a; for (int k = 0; k < K; ++k) r.v[k] = f.v[k] - g.v[k];
  return r;
}
template <int K> inline JetK<K> operator*(const JetK<K>& f, const JetK<K>& g) {
  JetK<K> r; r.a = f.a * g.a; for (int k = 0; k < K; ++k) r.v[k] = f.a * g.v[k] + f.v[k] * g.a;
  return r;
}
template <int K> inline JetK<K> operator/(const JetK<K>& f, const JetK<K>& g) {
  JetK<K> r; r.a = f.a / g.a;
  const double ig = 1.0 / g.a;
  for (int k = 0; k < K; ++k) r.v[k] = (f.v[k] - r.a * g.v[k]) * ig;
  return r;
}
template <int K> inline JetK<K> sqrt(const JetK<K>& f) {
  JetK<K> r; r.a = std::sqrt(f.a); const double h = 1.0 / (2.0 * r.a);
  for (int k = 0; k < K; ++k) r.v[k] = f.v[k] * h;
  return r;
}
template <int K> inline JetK<K> cos(const JetK<K>& f) {
  JetK<K> r; r.a = std::cos(f.a); const double d = -std::sin(f.a);
  for (int k = 0; k < K; ++k) r.v[k] = f.v[k] * d;
  return r;
}
template <int K> inline JetK<K> sin(const JetK<K>& f) {
  JetK<K> r; r.a = std::sin(f.a); const double d = std::cos(f.a);
  for (int k = 0; k < K; ++k) r.v[k] = f.v[k] * d;
  return r;
}
template <int K> inline double scalar_of(const JetK<K>& j) { return j.a; }

template <class T>
void angle_axis_rotate_point(const T w[3], const T pt[3], T out[3]) {
  const T theta2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (scalar_of(theta2) > std::numeric_limits<double>::epsilon()) {
    using std::sqrt; using std::cos; using std::sin;
    const T theta = sqrt(theta2);
    const T c = cos(theta), s = sin(theta);
    const T one(1.0);
    const T ti = one / theta;
    const T u[3] = {w[0] * ti, w[1] * ti, w[2] * ti};
    const T uxp[3] = {u[1] * pt[2] - u[2] * pt[1], u[2] * pt[0] - u[0] * pt[2], u[0] * pt[1] - u[1] * pt[0]};
    const T tmp = (u[0] * pt[0] + u[1] * pt[1] + u[2] * pt[2]) * (one - c);
    for (int k = 0; k < 3; ++k) out[k] = pt[k] * c + uxp[k] * s + u[k] * tmp;
  } else {
    const T wxp[3] = {w[1] * pt[2] - w[2] * pt[1], w[2] * pt[0] - w[0] * pt[2], w[0] * pt[1] - w[1] * pt[0]};
    for (int k = 0; k < 3; ++k) out[k] = pt[k] + wxp[k];
  }
}

// EdgeLossWithRot::operator() on (p1, p2, rot1, rot2)
template <class T>
void edge_rot_residual(const T* p1, const T* p2, const T* r1, const T* r2, const double* v, double lambda,
                       T* res) {
  T p[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  T q[3];
  angle_axis_rotate_point(r1, p, q);
  for (int k = 0; k < 3; ++k) res[k] = (q[k] - T(v[k])) * T(lambda);
  for (int k = 0; k < 3; ++k) res[3 + k] = (r1[k] - r2[k]) * T(1.0);
}

struct RotProblem {
  const double* grid; int N; int nv;
  std::vector<int> ha, hb;        // half-edges (a != b) in residual order
  std::vector<double> rest;       // v per half-edge
  double lambda;
};

// x: [6 nv] (p_i at 6i..6i+2, w_i at 6i+3..6i+5)
double rot_cost(const RotProblem& P, const double* x, double* vcost, double* ecost) {
  double cv = 0.0, ce = 0.0;
  for (int u = 0; u < P.nv; ++u) {
    const double* p = x + 6 * u;
    double d = distance_double_t<double>(P.grid, P.N, p, p[0], p[1], p[2]);
    cv += 0.5 * d * d;
  }
  for (size_t h = 0; h < P.ha.size(); ++h) {
    const int a = P.ha[h], b = P.hb[h];
    double r[6];
    edge_rot_residual<double>(x + 6 * a, x + 6 * b, x + 6 * a + 3, x + 6 * b + 3, &P.rest[3 * h], P.lambda, r);
    double s2 = 0.0;
    for (int k = 0; k < 6; ++k) s2 += r[k] * r[k];
    ce += 0.5 * s2;
  }
  if (vcost) *vcost = cv;
  if (ecost) *ecost = ce;
  return cv + ce;
}

// Per-block Jacobians at x: distance value d and gradient g (3 per vertex);
// edge residuals (6 per half-edge) and their 6x12 Jacobians over
// (p_a, p_b, w_a, w_b).
void rot_jacobian(const RotProblem& P, const double* x, std::vector<double>& d, std::vector<double>& g,
                  std::vector<double>& er, std::vector<double>& ej) {
  for (int u = 0; u < P.nv; ++u) {
    const double* p = x + 6 * u;
    Jet<double> pj[3] = {Jet<double>(p[0], 0), Jet<double>(p[1], 1), Jet<double>(p[2], 2)};
    Jet<double> r = distance_double_t<Jet<double>>(P.grid, P.N, pj, p[0], p[1], p[2]);
    d[u] = r.a;
    for (int k = 0; k < 3; ++k) g[3 * u + k] = r.v[k];
  }
  for (size_t h = 0; h < P.ha.size(); ++h) {
    const int a = P.ha[h], b = P.hb[h];
    JetK<12> p1[3], p2[3], r1[3], r2[3], res[6];
    for (int k = 0; k < 3; ++k) {
      p1[k] = JetK<12>(x[6 * a + k], k);
      p2[k] = JetK<12>(x[6 * b + k], 3 + k);
      r1[k] = JetK<12>(x[6 * a + 3 + k], 6 + k);
      r2[k] = JetK<12>(x[6 * b + 3 + k], 9 + k);
    }
    edge_rot_residual<JetK<12>>(p1, p2, r1, r2, &P.rest[3 * h], P.lambda, res);
    for (int r = 0; r < 6; ++r) {
      er[6 * h + r] = res[r].a;
      for (int c = 0; c < 12; ++c) ej[72 * h + 12 * r + c] = res[r].v[c];
    }
  }
}

// unknown index of column c (0..11) of a half-edge block: (vertex, local 0..5)
inline void rot_col(int c, int a, int b, int* v, int* l) {
  if (c < 3) { *v = a; *l = c; }
  else if (c < 6) { *v = b; *l = c - 3; }
  else if (c < 9) { *v = a; *l = 3 + (c - 6); }
  else { *v = b; *l = 3 + (c - 9); }
}

}  // namespace

// x: [nv,3] FP64 positions in the target's normalized frame (start and rest
// shape, updated in place); w: [nv,3] rotations (in/out; zeros as in the
// reference).  F: [nf,3].
OR_API int mo_oracle_lm_rot_solve(const double* grid, int N, double* xp, double* w, int nv, const int* F, int nf,
                                  double lambda, const OracleLMOptions* opt, OracleLMStats* st) {
  RotProblem P;
  P.grid = grid; P.N = N; P.nv = nv; P.lambda = lambda;
  for (int f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      const int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3];
      if (a == b) continue;
      P.ha.push_back(a);
      P.hb.push_back(b);
      for (int k = 0; k < 3; ++k) P.rest.push_back(xp[3 * a + k] - xp[3 * b + k]);
    }
  const int nh = (int)P.ha.size();
  const int n = 6 * nv;
  std::vector<double> x(n);
  for (int u = 0; u < nv; ++u)
    for (int k = 0; k < 3; ++k) { x[6 * u + k] = xp[3 * u + k]; x[6 * u + 3 + k] = w[3 * u + k]; }
  // vertex adjacency for the ordering and the envelope
  std::vector<std::vector<int>> adj(nv);
  for (int h = 0; h < nh; ++h) { adj[P.ha[h]].push_back(P.hb[h]); adj[P.hb[h]].push_back(P.ha[h]); }
  std::vector<int> rp(nv + 1, 0), ci;
  for (int u = 0; u < nv; ++u) {
    std::sort(adj[u].begin(), adj[u].end());
    adj[u].erase(std::unique(adj[u].begin(), adj[u].end()), adj[u].end());
    for (int v : adj[u]) ci.push_back(v);
    rp[u + 1] = (int)ci.size();
  }
  std::vector<int> perm = rcm_order(nv, rp, ci);
  Skyline S;
  S.n = n;
  S.first.resize(n);
  S.start.resize(n + 1);
  for (int u = 0; u < nv; ++u) {
    int pu = perm[u], mn = pu;
    for (int k = rp[u]; k < rp[u + 1]; ++k) mn = std::min(mn, perm[ci[k]]);
    for (int k = 0; k < 6; ++k) S.first[6 * pu + k] = 6 * mn;
  }
  size_t tot = 0;
  for (int i = 0; i < n; ++i) { S.start[i] = tot; tot += (size_t)(i - S.first[i] + 1); }
  S.start[n] = tot;
  S.vals.assign(tot, 0.0);

  std::vector<double> d(nv), g(3 * (size_t)nv), er(6 * (size_t)nh), ej(72 * (size_t)nh);
  std::vector<double> grad(n), cn(n), scale(n), diag(n), rhs(n), delta(n), cand(n);
  auto gradient_and_norms = [&]() {
    std::fill(grad.begin(), grad.end(), 0.0);
    std::fill(cn.begin(), cn.end(), 0.0);
    for (int u = 0; u < nv; ++u)
      for (int k = 0; k < 3; ++k) {
        grad[6 * u + k] += d[u] * g[3 * u + k];
        cn[6 * u + k] += g[3 * u + k] * g[3 * u + k];
      }
    for (int h = 0; h < nh; ++h)
      for (int c = 0; c < 12; ++c) {
        int v, l;
        rot_col(c, P.ha[h], P.hb[h], &v, &l);
        double sg = 0.0, sn = 0.0;
        for (int r = 0; r < 6; ++r) {
          const double J = ej[72 * h + 12 * r + c];
          sg += J * er[6 * h + r];
          sn += J * J;
        }
        grad[6 * v + l] += sg;
        cn[6 * v + l] += sn;
      }
  };
  double cost = rot_cost(P, x.data(), nullptr, nullptr);
  st->initial_cost = cost;
  rot_jacobian(P, x.data(), d, g, er, ej);
  gradient_and_norms();
  for (int c = 0; c < n; ++c) scale[c] = 1.0 / (1.0 + std::sqrt(cn[c]));

  double radius = opt->initial_trust_region_radius, decrease = 2.0;
  bool reuse = false, last_success = true;
  int iter = 0, invalid = 0, nsucc = 0, nfail = 0, term = 1;
  auto gmax = [&]() { double m = 0; for (int c = 0; c < n; ++c) m = std::max(m, std::fabs(grad[c])); return m; };
  auto idx = [&](int v, int l) { return 6 * perm[v] + l; };
  while (true) {
    if (iter >= opt->max_num_iterations) { term = 1; break; }
    if (last_success && gmax() <= opt->gradient_tolerance) { term = 0; break; }
    if (radius <= opt->min_trust_region_radius) { term = 0; break; }
    ++iter;
    if (!reuse)
      for (int c = 0; c < n; ++c)
        diag[c] = std::min(std::max(scale[c] * scale[c] * cn[c], opt->min_lm_diagonal), opt->max_lm_diagonal);
    reuse = true;
    // assemble S J^T J S + D^2 / radius (lower triangle, permuted order)
    std::fill(S.vals.begin(), S.vals.end(), 0.0);
    auto add = [&](int v1, int l1, int v2, int l2, double val) {
      int r = idx(v1, l1), c = idx(v2, l2);
      const double sv = scale[6 * v1 + l1] * scale[6 * v2 + l2] * val;
      if (c > r) return;            // each unordered pair is added from both orders
      S.at(r, c) += sv;
    };
    for (int u = 0; u < nv; ++u)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) add(u, k, u, l, g[3 * u + k] * g[3 * u + l]);
    for (int h = 0; h < nh; ++h)
      for (int c1 = 0; c1 < 12; ++c1) {
        int v1, l1;
        rot_col(c1, P.ha[h], P.hb[h], &v1, &l1);
        for (int c2 = 0; c2 < 12; ++c2) {
          int v2, l2;
          rot_col(c2, P.ha[h], P.hb[h], &v2, &l2);
          double sum = 0.0;
          for (int r = 0; r < 6; ++r) sum += ej[72 * h + 12 * r + c1] * ej[72 * h + 12 * r + c2];
          add(v1, l1, v2, l2, sum);
        }
      }
    for (int u = 0; u < nv; ++u)
      for (int l = 0; l < 6; ++l) S.at(idx(u, l), idx(u, l)) += diag[6 * u + l] / radius;
    bool ok = S.factor();
    double mcc = -1.0;
    if (ok) {
      for (int u = 0; u < nv; ++u)
        for (int l = 0; l < 6; ++l) rhs[idx(u, l)] = scale[6 * u + l] * grad[6 * u + l];
      S.solve(rhs);
      for (int u = 0; u < nv; ++u)
        for (int l = 0; l < 6; ++l) delta[6 * u + l] = scale[6 * u + l] * (-rhs[idx(u, l)]);
      for (int c = 0; c < n; ++c) if (!std::isfinite(delta[c])) ok = false;
    }
    if (ok) {
      double acc = 0.0;
      for (int u = 0; u < nv; ++u) {
        double m = 0.0;
        for (int k = 0; k < 3; ++k) m += g[3 * u + k] * delta[6 * u + k];
        acc += m * (d[u] + 0.5 * m);
      }
      for (int h = 0; h < nh; ++h)
        for (int r = 0; r < 6; ++r) {
          double m = 0.0;
          for (int c = 0; c < 12; ++c) {
            int v, l;
            rot_col(c, P.ha[h], P.hb[h], &v, &l);
            m += ej[72 * h + 12 * r + c] * delta[6 * v + l];
          }
          acc += m * (er[6 * h + r] + 0.5 * m);
        }
      mcc = -acc;
    }
    if (!ok || !(mcc > 0.0)) {
      ++invalid;
      if (invalid >= opt->max_consecutive_invalid_steps) { term = 2; break; }
      radius /= decrease; decrease *= 2.0;
      last_success = false; ++nfail;
      continue;
    }
    invalid = 0;
    for (int c = 0; c < n; ++c) cand[c] = x[c] + delta[c];
    double ccost = rot_cost(P, cand.data(), nullptr, nullptr);
    double sn = 0.0, xn = 0.0;
    for (int c = 0; c < n; ++c) { double t = x[c] - cand[c]; sn += t * t; xn += x[c] * x[c]; }
    sn = std::sqrt(sn); xn = std::sqrt(xn);
    double rho = (cost - ccost) / mcc;
    if (sn <= opt->parameter_tolerance * (xn + opt->parameter_tolerance)) { term = 0; break; }
    if (std::fabs(cost - ccost) <= opt->function_tolerance * cost) { term = 0; break; }
    if (rho > opt->min_relative_decrease) {
      x = cand;
      cost = ccost;
      rot_jacobian(P, x.data(), d, g, er, ej);
      gradient_and_norms();
      radius = radius / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3));
      radius = std::min(opt->max_trust_region_radius, radius);
      decrease = 2.0; reuse = false; last_success = true; ++nsucc;
    } else {
      radius /= decrease; decrease *= 2.0; reuse = true; last_success = false; ++nfail;
    }
  }
  double vc, ec;
  st->final_cost = rot_cost(P, x.data(), &vc, &ec);
  st->vertex_cost = vc; st->rigidity_cost = ec;
  st->iterations = iter; st->successful_steps = nsucc; st->unsuccessful_steps = nfail;
  st->termination = term;
  for (int u = 0; u < nv; ++u)
    for (int k = 0; k < 3; ++k) { xp[3 * u + k] = x[6 * u + k]; w[3 * u + k] = x[6 * u + 3 + k]; }
  return 0;
}

// J^T f of DeformWithRot at (x, w) with rest offsets from x0: grad [nv,6]
// (d/dp then d/dw per vertex), from the JetK Jacobians.
OR_API void mo_oracle_lm_rot_gradient(const double* grid, int N, const double* xp, const double* w,
                                      const double* x0, int nv, const int* F, int nf, double lambda, double* grad) {
  RotProblem P;
  P.grid = grid; P.N = N; P.nv = nv; P.lambda = lambda;
  for (int f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      const int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3];
      if (a == b) continue;
      P.ha.push_back(a);
      P.hb.push_back(b);
      for (int k = 0; k < 3; ++k) P.rest.push_back(x0[3 * a + k] - x0[3 * b + k]);
    }
  const int nh = (int)P.ha.size();
  std::vector<double> x(6 * (size_t)nv);
  for (int u = 0; u < nv; ++u)
    for (int k = 0; k < 3; ++k) { x[6 * u + k] = xp[3 * u + k]; x[6 * u + 3 + k] = w[3 * u + k]; }
  std::vector<double> d(nv), g(3 * (size_t)nv), er(6 * (size_t)nh), ej(72 * (size_t)nh);
  rot_jacobian(P, x.data(), d, g, er, ej);
  for (int c = 0; c < 6 * nv; ++c) grad[c] = 0.0;
  for (int u = 0; u < nv; ++u)
    for (int k = 0; k < 3; ++k) grad[6 * u + k] += d[u] * g[3 * u + k];
  for (int h = 0; h < nh; ++h)
    for (int c = 0; c < 12; ++c) {
      int v, l;
      rot_col(c, P.ha[h], P.hb[h], &v, &l);
      for (int r = 0; r < 6; ++r) grad[6 * v + l] += ej[72 * h + 12 * r + c] * er[6 * h + r];
    }
}

// DeformWithRot's final cost terms at (x, w) with rest offsets from x0.
OR_API double mo_oracle_lm_rot_cost(const double* grid, int N, const double* xp, const double* w, const double* x0,
                                    int nv, const int* F, int nf, double lambda, double* vcost, double* ecost) {
  RotProblem P;
  P.grid = grid; P.N = N; P.nv = nv; P.lambda = lambda;
  for (int f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      const int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3];
      if (a == b) continue;
      P.ha.push_back(a);
      P.hb.push_back(b);
      for (int k = 0; k < 3; ++k) P.rest.push_back(x0[3 * a + k] - x0[3 * b + k]);
    }
  std::vector<double> x(6 * (size_t)nv);
  for (int u = 0; u < nv; ++u)
    for (int k = 0; k < 3; ++k) { x[6 * u + k] = xp[3 * u + k]; x[6 * u + 3 + k] = w[3 * u + k]; }
  return rot_cost(P, x.data(), vcost, ecost);
}

// ---------------------------------------------------------------------------
// Grid build with an AABB tree over the faces, the way libigl's
// point_mesh_squared_distance queries its AABB (mesh.cc:140).  Same per-face
// arithmetic and tie rule as mo_oracle_build_grid, so the outputs are
// identical; used as the CPU baseline's grid build.
namespace {

struct BvhNode {
  double lo[3], hi[3];
  int left, right;    // children (internal) or -1
  int first, count;   // face range in the permuted order (leaf)
};

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<int> perm;
};

int bvh_build(Bvh& B, const std::vector<FaceData>& faces, const std::vector<V3>& cen, int first, int count) {
  BvhNode nd;
  for (int a = 0; a < 3; ++a) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
  double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = first; i < first + count; ++i) {
    const FaceData& f = faces[B.perm[i]];
    for (int j = 0; j < 3; ++j) {
      const double p[3] = {f.p[j].x, f.p[j].y, f.p[j].z};
      for (int a = 0; a < 3; ++a) { nd.lo[a] = std::min(nd.lo[a], p[a]); nd.hi[a] = std::max(nd.hi[a], p[a]); }
    }
    const double c[3] = {cen[B.perm[i]].x, cen[B.perm[i]].y, cen[B.perm[i]].z};
    for (int a = 0; a < 3; ++a) { clo[a] = std::min(clo[a], c[a]); chi[a] = std::max(chi[a], c[a]); }
  }
  nd.left = nd.right = -1;
  nd.first = first;
  nd.count = count;
  int id = (int)B.nodes.size();
  B.nodes.push_back(nd);
  if (count <= 4) return id;
  int ax = 0;
  for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
  int mid = first + count / 2;
  std::nth_element(B.perm.begin() + first, B.perm.begin() + mid, B.perm.begin() + first + count, [&](int x, int y) {
    const double cx = ax == 0 ? cen[x].x : (ax == 1 ? cen[x].y : cen[x].z);
    const double cy = ax == 0 ? cen[y].x : (ax == 1 ? cen[y].y : cen[y].z);
    return cx < cy || (cx == cy && x < y);
  });
  int l = bvh_build(B, faces, cen, first, mid - first);
  int r = bvh_build(B, faces, cen, mid, first + count - mid);
  B.nodes[id].left = l;
  B.nodes[id].right = r;
  return id;
}

inline double box_d2(const BvhNode& n, V3 q) {
  const double p[3] = {q.x, q.y, q.z};
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    double g = 0.0;
    if (p[a] < n.lo[a]) g = n.lo[a] - p[a];
    else if (p[a] > n.hi[a]) g = p[a] - n.hi[a];
    s += g * g;
  }
  return s;
}

}  // namespace

OR_API void mo_oracle_build_grid_bvh(const double* V, const int* F, long nf, int N, double* dist, float* dist32,
                                     int* nearest) {
  std::vector<FaceData> faces(nf);
  std::vector<V3> cen(nf);
  for (long f = 0; f < nf; ++f) {
    faces[f] = make_face(V, F, f);
    cen[f] = V3{(faces[f].p[0].x + faces[f].p[1].x + faces[f].p[2].x) / 3.0,
                (faces[f].p[0].y + faces[f].p[1].y + faces[f].p[2].y) / 3.0,
                (faces[f].p[0].z + faces[f].p[1].z + faces[f].p[2].z) / 3.0};
  }
  Bvh B;
  B.perm.resize(nf);
  for (long f = 0; f < nf; ++f) B.perm[f] = (int)f;
  if (nf > 0) bvh_build(B, faces, cen, 0, (int)nf);
  std::vector<int> stack;
  const long total = (long)N * N * N;
  for (long c = 0; c < total; ++c) {
    int k = (int)(c % N), j = (int)((c / N) % N), i = (int)(c / ((long)N * N));
    V3 q{(double)k / N, (double)j / N, (double)i / N};
    double best = INFINITY;
    int bi = -1;
    stack.clear();
    if (nf > 0) stack.push_back(0);
    while (!stack.empty()) {
      int id = stack.back();
      stack.pop_back();
      const BvhNode& n = B.nodes[id];
      if (box_d2(n, q) > best * (1.0 + 1e-9)) continue;
      if (n.left < 0) {
        for (int t = n.first; t < n.first + n.count; ++t) {
          int f = B.perm[t];
          double d = point_face_d2(faces[f], q);
          if (d < best || (d == best && f < bi)) { best = d; bi = f; }
        }
      } else {
        double dl = box_d2(B.nodes[n.left], q), dr = box_d2(B.nodes[n.right], q);
        if (dl <= dr) { stack.push_back(n.right); stack.push_back(n.left); }
        else { stack.push_back(n.left); stack.push_back(n.right); }
      }
    }
    double s = std::sqrt(best);
    if (dist) dist[c] = s;
    if (dist32) dist32[c] = (float)s;
    if (nearest) nearest[c] = bi;
  }
}
