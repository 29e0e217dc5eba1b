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
// Incidence lists for the backward gather, built on the device from the
// index buffer: vertex v's entries rp[v]..rp[v+1] are (h << 1) | role in
// increasing h, role 0 (tail, subtract) before role 1 (head, add) -- the
// order in which the reference's sequential loop updates that vertex.  Ends
// outside [0, nv) set *err and are left out.  ws holds
// edge_incidence_workspace_bytes(nh, nv) bytes; *rp and *inc point into it.
size_t edge_incidence_workspace_bytes(long long nh, long long nv);
int edge_incidence_build(const int* idx, int kind, long long nh, long long nv, void* ws, long long** rp,
                         long long** inc, int* err, hipStream_t s);
int edge_bwd_launch(const float* V, long long nv, const int* idx, int kind, const float* off, const long long* rp,
                    const long long* inc, float* out, hipStream_t s);

}  // namespace mo
