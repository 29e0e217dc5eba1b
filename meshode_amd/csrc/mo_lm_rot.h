// Levenberg-Marquardt solve of Deformer::DeformWithRot (deformer.cc:94-171)
// on gfx950: positions and per-vertex angle-axis rotations.
#pragma once

#include "mo_common.h"
#include "mo_lm.h"

namespace mo {

struct RotPairDesc {
  const float* grid;   // N^3 FP32 distance grid of the target
  int N;
  int nv;              // vertices
  int nh;              // half-edges (face edges with distinct endpoints)
  long long vbase;     // offset in the per-vertex arrays
  long long hbase;     // offset in the per-half-edge arrays
  long long lbase;     // offset in the vertex half-edge lists
  float lambda;        // EdgeLossWithRot lambda
};

constexpr int kRotVSlots = 28;   // float4 work slots per vertex
constexpr int kRotHSlots = 7;    // float4 work slots per half-edge

struct RotArrays {
  const float4* x0;    // rest position (normalized frame)
  float4* u;           // displacement from x0 (in/out)
  float4* w;           // angle-axis rotation (in/out)
  const int4* vadj;    // (first list entry, outgoing count, incoming count, 0)
  const int* vlist;    // local half-edge ids: the outgoing ones, then the incoming ones
  const int2* hab;     // local (a, b) of each half-edge, a = F[f][j], b = F[f][j+1]
  float4* vwork;       // kRotVSlots per vertex, slot-major over all vertices
  float4* hwork;       // kRotHSlots per half-edge, slot-major over all half-edges
};

int lm_rot_launch(const RotPairDesc* pairs_dev, int npairs, const RotArrays& arrays, const LMOptionsDev& opt,
                  LMResultDev* results_dev, long long total_vertices, long long total_half_edges, int block_threads,
                  hipStream_t s);

// Multi-workgroup variant (k_lm_rot_team), laid out as lm_team_launch
// (mo_lm.h): wg_dev entries (team, rank, size, -), team t solves problem
// plist[t]; cnt / part / abort_flag as there.
int lm_rot_team_capacity(int block_threads);
int lm_rot_team_launch(const RotPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg,
                       const RotArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                       long long total_vertices, long long total_half_edges, int block_threads, unsigned* cnt,
                       double* part, int* abort_flag, hipStream_t s);

}  // namespace mo
