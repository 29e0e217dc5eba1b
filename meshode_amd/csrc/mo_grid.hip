// Distance-field build: Mesh::ConstructDistanceField (mesh.cc:106-152) on
// gfx950, plus the template normalization (Mesh::Normalize, mesh.cc:66-96)
// and Mesh::ApplyTransform (mesh.cc:98-105) it depends on.
//
// Grid point (i, j, k) sits at (k/N, j/N, i/N) (mesh.cc:110-116) and its
// value is sqrt of the squared distance to the nearest target triangle
// (mesh.cc:140-150, igl::point_mesh_squared_distance).  The squared distance
// is computed in FP64 with a canonical operation sequence (face region when
// the projection falls inside, else the three edges with endpoints ordered
// by vertex index) so that triangles sharing an edge or vertex produce the
// identical value and ties resolve to the lowest triangle index.  This file
// is compiled with -ffp-contract=off: every FP64 operation is one IEEE
// rounding, so results are bit-reproducible on the host.
//
// Kernel structure (HBM traffic is tiny; the build is FP64-VALU bound):
//   1. k_face_bins   : per face, index check + Morton bin of the centroid
//                      (16^3 bins).
//   2. k_bin_sort    : one workgroup counting-sorts face indices by bin (LDS
//                      histogram + scan) so tiles of consecutive faces are
//                      spatially tight.
//   3. k_face_prep   : FP64 face record per face, written in sorted order.
//   4. k_tile_bounds : AABB per tile of kTile (8) sorted faces;
//      k_super_bounds: AABB per super-tile of kSuperTiles (16) tiles.
//   5. k_grid_build  : one wavefront per 4x4x8 brick of grid points (two
//                      points per lane, 4 apart in z).  The
//                      wave first evaluates the 6 tiles nearest its brick's
//                      centre among the tiles of the 4 nearest super-tiles
//                      (they seed every lane with a near-exact distance),
//                      then walks the super-tiles 64 at a time: a
//                      brick-level ballot drops super-tiles whose AABB is
//                      farther from the brick than the wave's worst best
//                      distance, the tiles of the kept ones (64 per pass)
//                      go through the same test, and each remaining tile is
//                      kept only if some lane's own point-to-AABB bound
//                      still reaches that lane's best.  For a kept tile,
//                      one lane per face first drops the faces whose AABB
//                      is out of reach of the whole brick; the survivors'
//                      records are copied into LDS with coalesced loads,
//                      and the wave walks them (one LDS address for all
//                      lanes), testing a face exactly only when its AABB
//                      and plane bounds can still improve some lane.  Ties
//                      keep the lowest face index, so the visiting order
//                      never changes the result.
#include "mo_common.h"
#include "mo_grid.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mo {

namespace {

// faces per culling tile: with the super-tile level below, 8 against 16
// gave 41.4 against 42.7 ms for 512 box-union grids at 64^3 (65 against
// 90 bound checks per cell: tighter tile AABBs) and the same time for the
// mixed 64^3 / 128^3 icosphere targets of up to 81,920 faces; 4 was
// slower (43.8 ms).  Without super-tiles 16 beat 32 and 8 (the scan over
// all tiles grew with their number).
constexpr int kTile = 8;
constexpr int kSeedTiles = 6;  // nearest tiles evaluated before the culled walk (4: 39.7, 8: 44.0 ms against 38.8)
// tiles per super-tile: the walk culls 128-face super-tiles against the
// brick first, then the tiles of the kept ones (four super-tiles, 64 tiles,
// per wavefront pass), so the walk's cost grows with the super-tiles near
// the brick rather than with the whole target (32 tiles per super-tile was
// slower: 47.7 ms, the seeds then come from two super-tiles)
constexpr int kSuperTiles = 16;
constexpr int kSupPass = 64 / kSuperTiles;   // super-tiles whose tiles one wavefront pass covers
static_assert(kSuperTiles * kSupPass == 64 && kSuperTiles <= 64, "a pass covers 64 tiles");
constexpr int kBinBits = 4;                  // 16 bins per axis
constexpr int kBins = 1 << (3 * kBinBits);   // 4096

__device__ inline unsigned spread3(unsigned v) {
  // spread the low 4 bits of v to every third bit
  unsigned r = 0;
  for (int b = 0; b < kBinBits; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

__device__ inline double dmin(double a, double b) { return b < a ? b : a; }
__device__ inline double dmax(double a, double b) { return b > a ? b : a; }
// One-instruction v_max_f64 / v_min_f64 for the culling bounds and the
// minimum of the three edge distances: the operands are never NaN there, and
// where two operands are equal (including +0 / -0) either result gives the
// same squared distance and the same comparisons.
__device__ inline double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ inline double vmin(double a, double b) { return __builtin_fmin(a, b); }

// Pass 1: validate indices and compute the centroid bin (Morton order of a
// 16^3 lattice over [0,1]^3, clamped).  blockIdx.y = job.
__global__ void k_face_bins(const GridJob* __restrict__ jobs, int* __restrict__ err) {
  const GridJob J = jobs[blockIdx.y];
  for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < J.nf;
       f += (long long)gridDim.x * blockDim.x) {
    double c[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j) {
      int vi = J.F[f * 3 + j];
      if (vi < 0 || vi >= J.nv) { atomicOr(err, 1); vi = 0; }
      for (int a = 0; a < 3; ++a) c[a] += J.V[(long long)vi * 3 + a];
    }
    unsigned b3[3];
    for (int a = 0; a < 3; ++a) {
      int b = (int)floor(c[a] * (1.0 / 3.0) * (1 << kBinBits));
      b = b < 0 ? 0 : (b > (1 << kBinBits) - 1 ? (1 << kBinBits) - 1 : b);
      b3[a] = (unsigned)b;
    }
    J.bin[f] = spread3(b3[0]) | (spread3(b3[1]) << 1) | (spread3(b3[2]) << 2);
  }
}

// Pass 2: one workgroup per job counting-sorts face indices by bin (LDS
// histogram, block-wide scan of the 4096 bins, LDS-atomic scatter).  The
// set-up kernels of a batch use workgroups of at most 256 threads (one
// wavefront per SIMD), which fit beside a running grid build's wavefronts,
// so a second batch's set-up is not held back behind them.
constexpr int kSortT = 256;
constexpr int kBinsPerThread = kBins / kSortT;   // 16 consecutive bins each
__global__ void __launch_bounds__(kSortT) k_bin_sort(const GridJob* __restrict__ jobs) {
  const GridJob J = jobs[blockIdx.x];
  __shared__ unsigned hist[kBins];
  __shared__ unsigned wsum[kSortT / 64];
  for (int b = threadIdx.x; b < kBins; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  for (long long f = threadIdx.x; f < J.nf; f += blockDim.x) atomicAdd(&hist[J.bin[f]], 1u);
  __syncthreads();
  // exclusive scan: each thread owns kBinsPerThread consecutive bins
  const int t = threadIdx.x;
  unsigned v[kBinsPerThread];
  unsigned mine = 0;
#pragma unroll
  for (int q = 0; q < kBinsPerThread; ++q) {
    v[q] = hist[kBinsPerThread * t + q];
    mine += v[q];
  }
  unsigned incl = mine;
  const int lane = t & 63, w = t >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    unsigned y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  unsigned base = 0;
  for (int q = 0; q < w; ++q) base += wsum[q];
  unsigned ex = base + incl - mine;
#pragma unroll
  for (int q = 0; q < kBinsPerThread; ++q) {
    hist[kBinsPerThread * t + q] = ex;
    ex += v[q];
  }
  __syncthreads();
  for (long long f = threadIdx.x; f < J.nf; f += blockDim.x) {
    unsigned pos = atomicAdd(&hist[J.bin[f]], 1u);
    J.perm[pos] = (int)f;
  }
}

// Pass 3: FP64 face record of face perm[pos], written at its sorted position.
// At most 64 VGPRs (8 wavefronts per SIMD): the kernel then fits beside a
// running grid build's wavefronts (85 VGPRs, 5 per SIMD), so a second
// batch's face preparation is not held back until that build ends.
__global__ void __launch_bounds__(256, 8) k_face_prep(const GridJob* __restrict__ jobs) {
  const GridJob J = jobs[blockIdx.y];
  const double* __restrict__ V = J.V;
  const int* __restrict__ F = J.F;
  const long long nv = J.nv;
  for (long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x; pos < J.nf;
       pos += (long long)gridDim.x * blockDim.x) {
  const long long f = J.perm[pos];
  GFace g;
  int idx[3];
  double p[3][3];
  for (int j = 0; j < 3; ++j) {
    int vi = F[f * 3 + j];
    if (vi < 0 || vi >= nv) vi = 0;     // reported by k_face_bins
    idx[j] = vi;
    p[j][0] = V[(long long)vi * 3 + 0];
    p[j][1] = V[(long long)vi * 3 + 1];
    p[j][2] = V[(long long)vi * 3 + 2];
  }
  g.ax = p[0][0]; g.ay = p[0][1]; g.az = p[0][2];
  g.bx = p[1][0]; g.by = p[1][1]; g.bz = p[1][2];
  g.cx = p[2][0]; g.cy = p[2][1]; g.cz = p[2][2];
  double e0x = g.bx - g.ax, e0y = g.by - g.ay, e0z = g.bz - g.az;
  double e1x = g.cx - g.ax, e1y = g.cy - g.ay, e1z = g.cz - g.az;
  g.nx = e0y * e1z - e0z * e1y;
  g.ny = e0z * e1x - e0x * e1z;
  g.nz = e0x * e1y - e0y * e1x;
  double nn = g.nx * g.nx + g.ny * g.ny + g.nz * g.nz;
  g.inv_nn = nn > 0.0 ? 1.0 / nn : 0.0;
  g.d00 = e0x * e0x + e0y * e0y + e0z * e0z;
  g.d01 = e0x * e1x + e0y * e1y + e0z * e1z;
  g.d11 = e1x * e1x + e1y * e1y + e1z * e1z;
  g.denom = g.d00 * g.d11 - g.d01 * g.d01;
  int flip = 0;
  double ll[3], il[3];
  for (int j = 0; j < 3; ++j) {
    int i0 = j, i1 = (j + 1) % 3;
    if (idx[i1] < idx[i0]) { flip |= 1 << j; int t = i0; i0 = i1; i1 = t; }
    double dx = p[i1][0] - p[i0][0], dy = p[i1][1] - p[i0][1], dz = p[i1][2] - p[i0][2];
    ll[j] = dx * dx + dy * dy + dz * dz;
    il[j] = ll[j] > 0.0 ? 1.0 / ll[j] : 0.0;
  }
  g.ll0 = ll[0]; g.ll1 = ll[1]; g.ll2 = ll[2];
  g.il0 = il[0]; g.il1 = il[1]; g.il2 = il[2];
  g.lox = dmin(g.ax, dmin(g.bx, g.cx)); g.hix = dmax(g.ax, dmax(g.bx, g.cx));
  g.loy = dmin(g.ay, dmin(g.by, g.cy)); g.hiy = dmax(g.ay, dmax(g.by, g.cy));
  g.loz = dmin(g.az, dmin(g.bz, g.cz)); g.hiz = dmax(g.az, dmax(g.bz, g.cz));
  g.flip = flip;
  g.face = (int)f;
  J.faces[pos] = g;
  J.cull[pos] = GCull{g.lox, g.loy, g.loz, g.hix, g.hiy, g.hiz, g.nx, g.ny, g.nz, g.ax, g.ay, g.az, g.inv_nn, 0.0};
  }
}

// One wavefront per tile: lanes read the face AABBs, wave min/max.
__global__ void __launch_bounds__(256) k_tile_bounds(const GridJob* __restrict__ jobs) {
  const GridJob J = jobs[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= J.ntiles) return;
  const long long f = t * kTile + lane;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (lane < kTile && f < J.nf) {
    const GCull* g = J.cull + f;
    v[0] = g->lox; v[1] = g->loy; v[2] = g->loz; v[3] = g->hix; v[4] = g->hiy; v[5] = g->hiz;
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int a = 0; a < 6; ++a) {
      double y = __shfl_xor(v[a], o, 64);
      v[a] = a < 3 ? dmin(v[a], y) : dmax(v[a], y);
    }
  if (lane < 6) {
    double out = v[0];
    for (int a = 1; a < 6; ++a) out = lane == a ? v[a] : out;
    J.tb[t * 6 + lane] = out;
  }
}

// One thread per super-tile: the union of its tiles' AABBs.
__global__ void k_super_bounds(const GridJob* __restrict__ jobs) {
  const GridJob J = jobs[blockIdx.y];
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= J.nsuper) return;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  const long long t1 = (s + 1) * kSuperTiles < J.ntiles ? (s + 1) * kSuperTiles : J.ntiles;
  for (long long t = s * kSuperTiles; t < t1; ++t)
    for (int a = 0; a < 3; ++a) {
      v[a] = dmin(v[a], J.tb[t * 6 + a]);
      v[3 + a] = dmax(v[3 + a], J.tb[t * 6 + 3 + a]);
    }
  for (int a = 0; a < 6; ++a) J.sb[s * 6 + a] = v[a];
}

// Canonical edge squared distance (endpoints ordered by vertex index).
__device__ inline double edge_d2(double qx, double qy, double qz, double ux, double uy, double uz,
                                 double wx, double wy, double wz, double ll, double il) {
  double dx = wx - ux, dy = wy - uy, dz = wz - uz;
  double px = qx - ux, py = qy - uy, pz = qz - uz;
  double tn = px * dx + py * dy + pz * dz;
  if (tn <= 0.0 || ll == 0.0) return px * px + py * py + pz * pz;
  if (tn >= ll) {
    double rx = qx - wx, ry = qy - wy, rz = qz - wz;
    return rx * rx + ry * ry + rz * rz;
  }
  double t = tn * il;
  double cx = ux + t * dx, cy = uy + t * dy, cz = uz + t * dz;
  double rx = qx - cx, ry = qy - cy, rz = qz - cz;
  return rx * rx + ry * ry + rz * rz;
}

__device__ inline double face_d2(const GFace& g, double qx, double qy, double qz) {
  double vx = qx - g.ax, vy = qy - g.ay, vz = qz - g.az;
  if (g.denom > 0.0 && g.inv_nn > 0.0) {
    double e0x = g.bx - g.ax, e0y = g.by - g.ay, e0z = g.bz - g.az;
    double e1x = g.cx - g.ax, e1y = g.cy - g.ay, e1z = g.cz - g.az;
    double d20 = vx * e0x + vy * e0y + vz * e0z;
    double d21 = vx * e1x + vy * e1y + vz * e1z;
    double bb = g.d11 * d20 - g.d01 * d21;
    double cc = g.d00 * d21 - g.d01 * d20;
    if (bb >= 0.0 && cc >= 0.0 && bb + cc <= g.denom) {
      double dp = g.nx * vx + g.ny * vy + g.nz * vz;
      return (dp * dp) * g.inv_nn;
    }
  }
  // edge j joins p_j and p_{j+1}; flip bit j set when p_{j+1} has the lower
  // index.  The face is the same in every lane, so the flip bits are read
  // into a scalar register and each edge takes one uniform branch.
  const int flip = __builtin_amdgcn_readfirstlane(g.flip);
  double d0, d1, d2;
  if (flip & 1) d0 = edge_d2(qx, qy, qz, g.bx, g.by, g.bz, g.ax, g.ay, g.az, g.ll0, g.il0);
  else d0 = edge_d2(qx, qy, qz, g.ax, g.ay, g.az, g.bx, g.by, g.bz, g.ll0, g.il0);
  if (flip & 2) d1 = edge_d2(qx, qy, qz, g.cx, g.cy, g.cz, g.bx, g.by, g.bz, g.ll1, g.il1);
  else d1 = edge_d2(qx, qy, qz, g.bx, g.by, g.bz, g.cx, g.cy, g.cz, g.ll1, g.il1);
  if (flip & 4) d2 = edge_d2(qx, qy, qz, g.ax, g.ay, g.az, g.cx, g.cy, g.cz, g.ll2, g.il2);
  else d2 = edge_d2(qx, qy, qz, g.cx, g.cy, g.cz, g.ax, g.ay, g.az, g.ll2, g.il2);
  return vmin(vmin(d2, d1), d0);
}

// Device memory reached through the job table's pointers: the explicit
// global address space makes the accesses global_* instead of flat_*
// instructions (a flat load also counts against the LDS wait counter, so
// every LDS wait would wait for it).
#if defined(__HIP_DEVICE_COMPILE__)
#define MO_GLOBAL __attribute__((address_space(1)))
#else
#define MO_GLOBAL
#endif
template <class T>
__device__ inline MO_GLOBAL T* gptr(T* p) {
  return (MO_GLOBAL T*)p;
}

// Wave-wide reductions through DPP lane moves (quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror, row_mirror), then the four row results by
// readlane: every lane gets the same value, with no LDS permutes.
template <int CTRL>
__device__ inline unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = dpp_u32<CTRL>((unsigned)b), hi = dpp_u32<CTRL>((unsigned)(b >> 32));
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

__device__ inline double wave_max(double v) {
  v = vmax(v, dpp_f64<0xB1>(v));
  v = vmax(v, dpp_f64<0x4E>(v));
  v = vmax(v, dpp_f64<0x141>(v));
  v = vmax(v, dpp_f64<0x140>(v));
  return vmax(vmax(readlane_f64(v, 0), readlane_f64(v, 16)), vmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// (value, index) with the smaller value, ties to the smaller non-negative
// index (-1: none)
__device__ inline void argmin_step(double& ol, int& ot, double xl, int xt) {
  if (xl < ol || (xl == ol && xt >= 0 && (ot < 0 || xt < ot))) { ol = xl; ot = xt; }
}
template <int CTRL>
__device__ inline void argmin_dpp(double& ol, int& ot) {
  argmin_step(ol, ot, dpp_f64<CTRL>(ol), (int)dpp_u32<CTRL>((unsigned)ot));
}
// the wave's argmin (the same in every lane)
__device__ inline void wave_argmin(double& ol, int& ot) {
  argmin_dpp<0xB1>(ol, ot);
  argmin_dpp<0x4E>(ol, ot);
  argmin_dpp<0x141>(ol, ot);
  argmin_dpp<0x140>(ol, ot);
  double rl = readlane_f64(ol, 0);
  int rt = __builtin_amdgcn_readlane(ot, 0);
  for (int r = 16; r < 64; r += 16) argmin_step(rl, rt, readlane_f64(ol, r), __builtin_amdgcn_readlane(ot, r));
  ol = rl;
  ot = rt;
}

__device__ inline double box_gap2(const MO_GLOBAL double* __restrict__ tb, long long t, double lx, double ly, double lz,
                                  double hx, double hy, double hz) {
  double tl[3] = {tb[t * 6 + 0], tb[t * 6 + 1], tb[t * 6 + 2]};
  double th[3] = {tb[t * 6 + 3], tb[t * 6 + 4], tb[t * 6 + 5]};
  double ql[3] = {lx, ly, lz}, qh[3] = {hx, hy, hz};
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    double g = vmax(0.0, vmax(tl[a] - qh[a], ql[a] - th[a]));
    s += g * g;
  }
  return s;
}

// One wavefront (64 lanes) per 4x4x8 brick: lane l holds the two grid
// points (x, y, z) and (x, y, z + 4) of it.  The two points share the
// brick's culling, the staged face records, every LDS read of a record and
// the x / y parts of their bounds.
constexpr int kZPts = 2;   // points per lane, 4 apart in z
__global__ void __launch_bounds__(64) k_grid_build(const GridJob* __restrict__ jobs) {
  // survivors of a tile's brick-level cull: their full records, copied with
  // coalesced vector loads, then read by the whole wave at one LDS address
  __shared__ GFace s_face[kTile];
  __shared__ int s_list[kTile];
  const GridJob& JJ = jobs[blockIdx.y];
  const MO_GLOBAL GFace* __restrict__ faces = gptr(JJ.faces);
  const MO_GLOBAL GCull* __restrict__ cull = gptr(JJ.cull);
  const long long nf = JJ.nf;
  const MO_GLOBAL double* __restrict__ tb = gptr(JJ.tb);
  const long long ntiles = JJ.ntiles;
  const MO_GLOBAL double* __restrict__ sb = gptr(JJ.sb);
  const long long nsuper = JJ.nsuper;
  const int N = JJ.N;
  const int bpa = (N + 3) / 4, bpz = (N + 4 * kZPts - 1) / (4 * kZPts);
  MO_GLOBAL double* __restrict__ dist64 = gptr(JJ.dist64);
  MO_GLOBAL float* __restrict__ dist32 = gptr(JJ.dist32);
  MO_GLOBAL int* __restrict__ nearest = gptr(JJ.nearest);
  const int lane = threadIdx.x;
  long long b = blockIdx.x;
  if (b >= (long long)bpa * bpa * bpz) return;
  const int bx = (int)(b % bpa);
  const int by = (int)((b / bpa) % bpa);
  const int bz = (int)(b / ((long long)bpa * bpa));
  const int k = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3);
  int ip[kZPts];
  bool valid[kZPts];
  double qz[kZPts];
  const double qx = (double)k / (double)N, qy = (double)j / (double)N;
#pragma unroll
  for (int s = 0; s < kZPts; ++s) {
    ip[s] = bz * 4 * kZPts + (lane >> 4) + 4 * s;
    valid[s] = (k < N) && (j < N) && (ip[s] < N);
    qz[s] = (double)ip[s] / (double)N;
  }
  // brick AABB over valid points
  const int k1 = min(bx * 4 + 3, N - 1), j1 = min(by * 4 + 3, N - 1), i1 = min(bz * 4 * kZPts + 4 * kZPts - 1, N - 1);
  const double lx = (double)(bx * 4) / N, ly = (double)(by * 4) / N, lz = (double)(bz * 4 * kZPts) / N;
  const double hx = (double)k1 / N, hy = (double)j1 / N, hz = (double)i1 / N;

  // points farther than the band only need to be shown farther (GridSpec)
  double best[kZPts], bcut[kZPts];   // bcut: the culling threshold, kept with best
  int bidx[kZPts];
#pragma unroll
  for (int s = 0; s < kZPts; ++s) {
    best[s] = JJ.band2;
    bcut[s] = best[s] * (1.0 + 1e-9);
    bidx[s] = INT_MAX;
  }
  // the wave's worst best over its valid points
  auto worst = [&]() {
    double m = -INFINITY;
#pragma unroll
    for (int s = 0; s < kZPts; ++s) m = valid[s] ? vmax(m, best[s]) : m;
    return wave_max(m);
  };
  // work counters (wave-uniform): exact evaluations (one point per lane),
  // lanes that needed one, face bound checks (per point), staged tiles
  unsigned long long n_exact = 0, n_needed = 0, n_bound = 0, n_tiles = 0;

  // bound: an upper bound of every valid point's best (the wave's worst
  // best when it was taken; bests only shrink, so an older value stays valid)
  auto process_tile = [&](int t, double bound) {
    const long long f0 = (long long)t * kTile;
    const int cnt = (int)((nf - f0) < kTile ? (nf - f0) : kTile);
    // stage 1, one face per lane: drop faces whose AABB is farther from the
    // brick AABB than the wave's worst best distance
    bool cand = false;
    if (lane < cnt) {
      const MO_GLOBAL GCull* gp = cull + f0 + lane;
      const double ax = vmax(0.0, vmax(gp->lox - hx, lx - gp->hix));
      const double ay = vmax(0.0, vmax(gp->loy - hy, ly - gp->hiy));
      const double az = vmax(0.0, vmax(gp->loz - hz, lz - gp->hiz));
      cand = !(ax * ax + ay * ay + az * az > bound * (1.0 + 1e-9));
    }
    const unsigned long long m = __ballot(cand);
    const int ns = __popcll(m);
    if (ns == 0) return;
    ++n_tiles;
    n_bound += (unsigned long long)ns * kZPts;
    // compact the survivors (in face order) and stage their records in LDS
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (cand) s_list[rank] = lane;
    __syncthreads();
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    constexpr int kChunks = (int)(sizeof(GFace) / sizeof(u32x4));
    static_assert(sizeof(GFace) % sizeof(u32x4) == 0, "face records are copied in 16-byte chunks");
    const MO_GLOBAL u32x4* __restrict__ src = reinterpret_cast<const MO_GLOBAL u32x4*>(faces + f0);
    u32x4* dst = reinterpret_cast<u32x4*>(s_face);
    for (int c = lane; c < ns * kChunks; c += 64) {
      const int r = c / kChunks, w = c - r * kChunks;
      dst[c] = src[s_list[r] * kChunks + w];
    }
    __syncthreads();
    // stage 2, per surviving face: every point's own AABB bound, then (when
    // it leaves some point) the plane bound, then the exact test when some
    // lane's point can still improve or tie its best.  The bounds are summed
    // in the order x, y, z as for a single point.
    for (int r = 0; r < ns; ++r) {
      const GFace& g = s_face[r];
      const double gx = vmax(0.0, vmax(g.lox - qx, qx - g.hix));
      const double gy = vmax(0.0, vmax(g.loy - qy, qy - g.hiy));
      const double gxy = gx * gx + gy * gy;
      bool mine[kZPts];
      bool any = false;
#pragma unroll
      for (int s = 0; s < kZPts; ++s) {
        const double gz = vmax(0.0, vmax(g.loz - qz[s], qz[s] - g.hiz));
        mine[s] = valid[s] && !(gxy + gz * gz > bcut[s]);
        any = any || mine[s];
      }
      if (!__ballot(any)) continue;
      // the face plane is a second lower bound (0 for degenerate faces)
      const double dxy = g.nx * (qx - g.ax) + g.ny * (qy - g.ay);
#pragma unroll
      for (int s = 0; s < kZPts; ++s) {
        const double dp = dxy + g.nz * (qz[s] - g.az);
        mine[s] = mine[s] && !((dp * dp) * g.inv_nn > bcut[s]);
      }
#pragma unroll
      for (int s = 0; s < kZPts; ++s) {
        const unsigned long long need = __ballot(mine[s]);
        if (!need) continue;
        ++n_exact;
        n_needed += (unsigned long long)__popcll(need);
        // only the lanes that need it evaluate (a point whose bound exceeds
        // its best cannot improve or tie it), so the face-region / edge
        // branches diverge over those lanes alone
        if (mine[s]) {
          const double d = face_d2(g, qx, qy, qz[s]);
          if (d < best[s] || (d == best[s] && g.face < bidx[s])) {
            best[s] = d;
            bcut[s] = d * (1.0 + 1e-9);
            bidx[s] = g.face;
          }
        }
      }
    }
    __syncthreads();  // the next tile overwrites s_face / s_list
  };

  // 1. seeds: the kSupPass super-tiles nearest to the brick's centre (one
  //    scan of the super-tile AABBs; each lane keeps its nearest), then the
  //    kSeedTiles nearest of their 64 tiles, are evaluated first.  They give
  //    every point a near-exact distance, which makes the culling below
  //    tight (the choice of seeds affects speed only).
  int seeds[kSeedTiles];
  int nseed = 0;
  {
    // ranked by the AABB gap from the brick's centre (a brick inside
    // several AABBs has gap 0 to all of them)
    const double cx = 0.5 * (lx + hx), cy = 0.5 * (ly + hy), cz = 0.5 * (lz + hz);
    double slb = INFINITY;
    int ss = -1;
    for (long long c = 0; c < nsuper; c += 64) {
      const long long sidx = c + lane;
      if (sidx >= nsuper) continue;
      const double lb = box_gap2(sb, sidx, cx, cy, cz, cx, cy, cz);
      if (lb < slb) { slb = lb; ss = (int)sidx; }
    }
    int sp[kSupPass];
    for (int u = 0; u < kSupPass; ++u) sp[u] = -1;
    for (int u = 0; u < kSupPass; ++u) {
      double ol = slb;
      int os = ss;
      wave_argmin(ol, os);
      if (os < 0) break;
      if (ss == os) { slb = INFINITY; ss = -1; }
      sp[u] = os;
    }
    const int q = lane / kSuperTiles;
    int sq = -1;
#pragma unroll
    for (int u = 0; u < kSupPass; ++u) sq = q == u ? sp[u] : sq;
    double mlb = INFINITY;
    int mt = -1;
    if (sq >= 0) {
      const long long t = (long long)sq * kSuperTiles + (lane & (kSuperTiles - 1));
      if (t < ntiles) {
        mlb = box_gap2(tb, t, cx, cy, cz, cx, cy, cz);
        mt = (int)t;
      }
    }
    for (int r = 0; r < kSeedTiles; ++r) {
      double ol = mlb;
      int ot = mt;
      wave_argmin(ol, ot);
      if (ot < 0) break;
      if (mt == ot) { mlb = INFINITY; mt = -1; }
      seeds[nseed++] = ot;
    }
  }
  for (int r = 0; r < nseed; ++r) process_tile(seeds[r], worst());

  // 2. every other tile that can still improve (or tie) some point's best:
  //    super-tiles whose AABB is within the wave's worst best of the brick
  //    AABB, then (kSupPass kept super-tiles per pass, lane kSuperTiles q + j
  //    on tile j of the q-th) their tiles by the same test, and each such
  //    tile is kept only if some point's own point-to-AABB bound reaches its
  //    own best.
  for (long long c = 0; c < nsuper; c += 64) {
    const long long sidx = c + lane;
    const double slb = sidx < nsuper ? box_gap2(sb, sidx, lx, ly, lz, hx, hy, hz) : INFINITY;
    const double sbound = worst();
    unsigned long long smask = __ballot(sidx < nsuper && !(slb > sbound * (1.0 + 1e-9)));
    while (smask) {
      int sp[kSupPass];
      for (int u = 0; u < kSupPass; ++u) {
        sp[u] = smask ? (int)c + __ffsll((long long)smask) - 1 : -1;
        smask &= smask - 1;
      }
      const int q = lane / kSuperTiles;
      int sq = -1;
#pragma unroll
      for (int u = 0; u < kSupPass; ++u) sq = q == u ? sp[u] : sq;
      const long long t = sq < 0 ? -1 : (long long)sq * kSuperTiles + (lane & (kSuperTiles - 1));
      const double bound = worst();
      bool need = false;
      if (t >= 0 && t < ntiles) {
        bool taken = false;
        for (int r = 0; r < nseed; ++r) taken |= (seeds[r] == (int)t);
        need = !taken && !(box_gap2(tb, t, lx, ly, lz, hx, hy, hz) > bound * (1.0 + 1e-9));
      }
      unsigned long long tmask = __ballot(need);
      while (tmask) {
        const int l = __ffsll((long long)tmask) - 1;
        tmask &= tmask - 1;
        const int tt = __builtin_amdgcn_readlane((int)t, l);
        bool reach = false;
#pragma unroll
        for (int s = 0; s < kZPts; ++s)
          reach = reach || (valid[s] && !(box_gap2(tb, tt, qx, qy, qz[s], qx, qy, qz[s]) > bcut[s]));
        if (!__ballot(reach)) continue;
        process_tile(tt, bound);
      }
    }
  }
  if (JJ.stats && lane == 0) {
    atomicAdd(JJ.stats + kGridStatExact, n_exact * 64ull);
    atomicAdd(JJ.stats + kGridStatExactNeeded, n_needed);
    atomicAdd(JJ.stats + kGridStatBound, n_bound * 64ull);
    atomicAdd(JJ.stats + kGridStatTiles, n_tiles);
  }
#pragma unroll
  for (int s = 0; s < kZPts; ++s) {
    if (!valid[s]) continue;
    const long long cell = ((long long)ip[s] * N + j) * N + k;
    const double sq = sqrt(best[s]);
    if (dist64) dist64[cell] = sq;
    if (dist32) dist32[cell] = (float)sq;
    if (nearest) nearest[cell] = bidx[s] == INT_MAX ? -1 : bidx[s];
  }
}

// ---------------------------------------------------------------------------
// Normalization (Mesh::Normalize, mesh.cc:66-96) on device, FP64.

__device__ inline double xf_read(const XformJob& J, long long i) {
  return J.f64 ? ((const double*)J.V)[i] : (double)((const float*)J.V)[i];
}

// One workgroup per job: min / max per axis, then scale and pos (exact
// whatever the workgroup size: minima and maxima do not round).
constexpr int kStatT = 256;
__global__ void __launch_bounds__(kStatT) k_normalize_stats(const XformJob* __restrict__ jobs) {
  const XformJob J = jobs[blockIdx.x];
  __shared__ double smn[3][kStatT], smx[3][kStatT];
  double mn[3] = {1e30, 1e30, 1e30}, mx[3] = {-1e30, -1e30, -1e30};
  for (long long i = threadIdx.x; i < J.n; i += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      double v = xf_read(J, i * 3 + a);
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  for (int a = 0; a < 3; ++a) { smn[a][threadIdx.x] = mn[a]; smx[a][threadIdx.x] = mx[a]; }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        double o = smn[a][threadIdx.x + s];
        if (o < smn[a][threadIdx.x]) smn[a][threadIdx.x] = o;
        o = smx[a][threadIdx.x + s];
        if (o > smx[a][threadIdx.x]) smx[a][threadIdx.x] = o;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double ex = smx[0][0] - smn[0][0], ey = smx[1][0] - smn[1][0], ez = smx[2][0] - smn[2][0];
    double m = ez > ey ? ez : ey;           // std::max(b, c)
    m = m > ex ? m : ex;                    // std::max(a, .)
    double scale = m * 1.1;
    J.sp[0] = scale;
    for (int a = 0; a < 3; ++a) J.sp[1 + a] = smn[a][0] - 0.05 * scale;
  }
}

// out = (V - pos) / scale with the job's (scale, pos); blockIdx.y = job.
__global__ void k_apply_transform(const XformJob* __restrict__ jobs) {
  const XformJob J = jobs[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < J.n * 3;
       i += (long long)gridDim.x * blockDim.x) {
    int a = (int)(i % 3);
    J.out[i] = (xf_read(J, i) - J.sp[1 + a]) / J.sp[0];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t grid_job_workspace_bytes(long long nf) {
  const long long ntiles = (nf + kTile - 1) / kTile;
  const long long nsuper = (ntiles + kSuperTiles - 1) / kSuperTiles;
  return al256((size_t)nf * sizeof(GFace)) + al256((size_t)nf * sizeof(GCull)) + al256((size_t)nf * sizeof(unsigned)) +
         al256((size_t)nf * sizeof(int)) + al256((size_t)ntiles * 6 * sizeof(double)) +
         al256((size_t)nsuper * 6 * sizeof(double));
}

int grid_build_batch_launch(const GridSpec* specs, int n, void* workspace, hipStream_t stream, int* err_dev,
                            hipEvent_t build_after) {
  if (n <= 0) return MO_OK;
  std::vector<GridJob> jobs(n);
  char* w = (char*)workspace;
  GridJob* d_jobs = (GridJob*)w;
  w += al256((size_t)n * sizeof(GridJob));
  long long max_nf = 0, max_tiles = 0, max_bricks = 0, max_super = 0;
  for (int q = 0; q < n; ++q) {
    const GridSpec& sp = specs[q];
    MO_CHECK(sp.N >= 2 && sp.N <= 2048, MO_ERR_ARG, "grid resolution %d out of range [2, 2048]", sp.N);
    MO_CHECK(sp.nf >= 1, MO_ERR_ARG, "distance field needs at least one triangle (got %lld)", sp.nf);
    GridJob& J = jobs[q];
    J.V = sp.Vn; J.F = sp.F; J.nv = sp.nv; J.nf = sp.nf; J.N = sp.N;
    J.ntiles = (sp.nf + kTile - 1) / kTile;
    J.faces = (GFace*)w; w += al256((size_t)sp.nf * sizeof(GFace));
    J.cull = (GCull*)w; w += al256((size_t)sp.nf * sizeof(GCull));
    J.bin = (unsigned*)w; w += al256((size_t)sp.nf * sizeof(unsigned));
    J.perm = (int*)w; w += al256((size_t)sp.nf * sizeof(int));
    J.tb = (double*)w; w += al256((size_t)J.ntiles * 6 * sizeof(double));
    J.nsuper = (J.ntiles + kSuperTiles - 1) / kSuperTiles;
    J.sb = (double*)w; w += al256((size_t)J.nsuper * 6 * sizeof(double));
    J.dist64 = sp.dist64; J.dist32 = sp.dist32; J.nearest = sp.nearest;
    MO_CHECK(sp.band2 > 0.0, MO_ERR_ARG, "grid band must be positive");
    J.band2 = sp.band2;
    J.stats = sp.stats;
    const long long bpa = (sp.N + 3) / 4, bpz = (sp.N + 4 * kZPts - 1) / (4 * kZPts);
    max_nf = std::max(max_nf, sp.nf);
    max_tiles = std::max(max_tiles, J.ntiles);
    max_super = std::max(max_super, J.nsuper);
    max_bricks = std::max(max_bricks, bpa * bpa * bpz);
  }
  int rc = chain_h2d(d_jobs, jobs.data(), (size_t)n * sizeof(GridJob), stream);
  if (rc) return rc;
  const int T = 256;
  const unsigned fb = (unsigned)std::min<long long>((max_nf + T - 1) / T, 1024);
  hipLaunchKernelGGL(k_face_bins, dim3(fb, n), dim3(T), 0, stream, d_jobs, err_dev);
  hipLaunchKernelGGL(k_bin_sort, dim3(n), dim3(kSortT), 0, stream, d_jobs);
  hipLaunchKernelGGL(k_face_prep, dim3(fb, n), dim3(T), 0, stream, d_jobs);
  hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)((max_tiles + 3) / 4), n), dim3(256), 0, stream, d_jobs);
  hipLaunchKernelGGL(k_super_bounds, dim3((unsigned)((max_super + 63) / 64), n), dim3(64), 0, stream, d_jobs);
  if (build_after) MO_HIP(hipStreamWaitEvent(stream, build_after, 0));
  hipLaunchKernelGGL(k_grid_build, dim3((unsigned)max_bricks, n), dim3(64), 0, stream, d_jobs);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

size_t grid_batch_workspace_bytes(const GridSpec* specs, int n) {
  size_t b = al256((size_t)n * sizeof(GridJob));
  for (int q = 0; q < n; ++q) b += grid_job_workspace_bytes(specs[q].nf);
  return b + 256;
}

size_t xform_workspace_bytes(int n) { return al256((size_t)n * sizeof(XformJob)) + 256; }

int xform_batch_launch(const XformJob* jobs, int n, bool compute_stats, void* workspace, hipStream_t stream) {
  if (n <= 0) return MO_OK;
  XformJob* d_jobs = (XformJob*)workspace;
  long long max_n = 0;
  for (int q = 0; q < n; ++q) max_n = std::max(max_n, jobs[q].n);
  int rc = chain_h2d(d_jobs, jobs, (size_t)n * sizeof(XformJob), stream);
  if (rc) return rc;
  if (compute_stats) hipLaunchKernelGGL(k_normalize_stats, dim3(n), dim3(kStatT), 0, stream, d_jobs);
  if (max_n > 0) {
    const unsigned bx = (unsigned)std::min<long long>((max_n * 3 + 255) / 256, 1024);
    hipLaunchKernelGGL(k_apply_transform, dim3(bx, n), dim3(256), 0, stream, d_jobs);
  }
  MO_HIP(hipGetLastError());
  return MO_OK;
}

}  // namespace mo
