// Distance-field build and normalization launchers.
#pragma once

#include "mo_common.h"

namespace mo {

// FP64 face record for the point-triangle distance (240 bytes).
struct GFace {
  double ax, ay, az, bx, by, bz, cx, cy, cz;   // positions in face order
  double nx, ny, nz, inv_nn;                   // n = (b-a) x (c-a), 1/(n.n)
  double d00, d01, d11, denom;                 // Gram entries of (b-a, c-a)
  double ll0, ll1, ll2, il0, il1, il2;         // |edge j|^2 and its reciprocal
  double lox, loy, loz, hix, hiy, hiz;         // face AABB (culling only)
  int flip;                                    // bit j: edge j endpoints swapped (index order)
  int face;                                    // original face index
};

// Compact culling record of a face (112 bytes): AABB and plane.
struct GCull {
  double lox, loy, loz, hix, hiy, hiz;
  double nx, ny, nz, ax, ay, az, inv_nn, pad;
};

// One distance grid to build: normalized FP64 target vertices [nv,3] and
// int32 faces [nf,3] (device); any of dist64 / dist32 / nearest may be null.
// band2: squared distance beyond which a point only needs to be known to be
// farther: it stores sqrt(min(d^2, band2)) and nearest -1 when no face is
// within the band (INFINITY: the exact field everywhere).
struct GridSpec {
  const double* Vn;
  long long nv;
  const int* F;
  long long nf;
  int N;
  double* dist64;
  float* dist32;
  int* nearest;
  double band2 = INFINITY;
  unsigned long long* stats = nullptr;   // optional work counters (kGridStat*), added to
};

// Work counters of a grid build (per launch, summed over its grids), added
// with one atomic per wavefront when GridSpec::stats is set: point-triangle
// distance evaluations executed (64 lanes per wavefront evaluation) and the
// ones some lane needed, face bound checks executed (64 lanes each), and
// tiles of faces staged into LDS.
enum GridStat { kGridStatExact = 0, kGridStatExactNeeded, kGridStatBound, kGridStatTiles, kGridStatCount };

// The band a template grid needs: the lookups (uniformgrid.cc:18-150)
// interpolate the 8 corners of a cell and return 0 when the result exceeds
// 0.2.  The field is 1-Lipschitz, so a cell with a corner at d <= 0.2 has all
// corners within 0.2 + sqrt(3)/N; a cell with every corner beyond that is
// truncated, and stays truncated when its corners hold any value above
// 0.2 + 0.01.  Values clamped at this band therefore give every lookup the
// same result, value and gradient, as the exact field.
inline double lookup_band2(int N) {
  const double b = 0.2 + 1.7320508075688772 / N + 0.01;
  return b * b;
}

// Device-side job record (workspace slices included).
struct GridJob {
  const double* V;
  const int* F;
  long long nv, nf, ntiles;
  int N;
  GFace* faces;
  GCull* cull;
  unsigned* bin;
  int* perm;
  double* tb;
  long long nsuper;   // super-tiles: kSuperTiles consecutive tiles each
  double* sb;         // their AABBs [nsuper][6]
  double* dist64;
  float* dist32;
  int* nearest;
  double band2;
  unsigned long long* stats;
};

size_t grid_batch_workspace_bytes(const GridSpec* specs, int n);

// Builds every grid of the batch with one launch per stage.  err_dev: device
// int, bit 0 set when a face index is out of range.
// build_after (optional): the grid build proper (not the face preparation
// before it) waits for this event on `stream`.
int grid_build_batch_launch(const GridSpec* specs, int n, void* workspace, hipStream_t stream, int* err_dev,
                            hipEvent_t build_after = nullptr);

// Mesh::Normalize / Mesh::ApplyTransform jobs: V [n,3] (FP32 or FP64),
// sp = (scale, pos[3]) in device memory, out = (V - pos) / scale (FP64).
struct XformJob {
  const void* V;
  int f64;
  long long n;
  double* sp;
  double* out;
};

size_t xform_workspace_bytes(int n);

// compute_stats: also compute sp from V first (Mesh::Normalize); otherwise
// sp is an input (Mesh::ApplyTransform).  out may be null when n == 0.
int xform_batch_launch(const XformJob* jobs, int n, bool compute_stats, void* workspace, hipStream_t stream);

}  // namespace mo
