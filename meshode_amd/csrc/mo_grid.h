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

int grid_workspace_bytes(long long nf, size_t* bytes);

// Vn: normalized target vertices [nv,3] FP64 (device); F: [nf,3] int32
// (device).  Any of dist64 / dist32 / nearest may be null.  err_dev: device
// int, bit 0 set when a face index is out of range.
int grid_build_launch(const double* Vn, long long nv, const int* F, long long nf, int N, double* dist64,
                      float* dist32, int* nearest, void* workspace, hipStream_t stream, int* err_dev);

// Mesh::Normalize: sp_dev <- (scale, pos[3]); Vn_dev (optional) <- normalized copy.
template <class T>
int normalize_launch(const T* V, long long nv, double* sp_dev, double* Vn_dev, hipStream_t stream);

// Mesh::ApplyTransform with device-resident (scale, pos).
template <class T>
int apply_transform_launch(const T* V, long long nv, const double* sp_dev, double* out, hipStream_t stream);

}  // namespace mo
