// pyDeform layer kernel launchers.
#pragma once

#include "mo_common.h"

namespace mo {

// (De)NormalizeByTemplate on an FP32 [nv,3] device buffer, in place.
int normalize_by_template_launch(float* V, long long nv, double scale, const double* trans, bool inverse,
                                 hipStream_t s);

// DistanceFieldLoss forward (out [nv]) or backward (out [nv,3]).
int distance_loss_launch(const float* grid, int N, const float* V, long long nv, float* out, bool backward,
                         hipStream_t s);

// Edge layouts: kind 0 = faces [F,3] (3F directed half-edges),
// kind 1 = explicit edges [E,2].
int edge_store_launch(const float* V, long long nv, const int* idx, int kind, long long nh, float* off, int* err,
                      hipStream_t s);
int edge_fwd_launch(const float* V, long long nv, const int* idx, int kind, long long nh, const float* off,
                    float* out, int* err, hipStream_t s);
int edge_bwd_launch(const float* V, long long nv, const int* idx, int kind, const float* off, const long long* rp,
                    const long long* inc, float* out, hipStream_t s);

}  // namespace mo
