// Batched Levenberg-Marquardt deformation solver (Deformer::Deform /
// Deformer::DeformGraph, deformer.cc:18-92 and 370-442) on gfx950.
#pragma once

#include <vector>

#include <mutex>

#include "mo_common.h"

namespace mo {

// One source/target pair of the batch.  All pointers are device pointers.
struct LMPairDesc {
  const float* grid;      // N^3 FP32 distance grid of the target
  int N;
  int nv;                 // source vertices
  long long vbase;        // offset of this pair in the per-vertex arrays
  long long ebase;        // offset of this pair in the adjacency entry array
  float lambda;           // rigidity weight (EdgeLoss lambda); 1 for weighted pairs
  int flags;              // bit 0: per-entry edge weights in LMArrays::wt
};

// Solver options; defaults restate Ceres' Solver::Options defaults that
// deformer.cc:55-59 leaves untouched (max_num_iterations = 100).
struct LMOptionsDev {
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
  int pcg_max_iterations;       // inner PCG iteration cap per LM step
  float pcg_relative_tolerance; // ||r|| <= tol * ||r0|| stops PCG
  int max_successful_steps;     // > 0: stop after that many accepted steps
};

struct LMResultDev {
  double initial_cost;
  double final_cost;
  double vertex_cost;
  double rigidity_cost;
  int iterations;
  int successful_steps;
  int unsuccessful_steps;
  int termination;      // 0 convergence, 1 iteration cap, 2 failure
  long long pcg_iterations;
  long long clock_start;  // wall_clock64() at the start / end of the pair's solve
  long long clock_end;
};

// Number of float4 scratch slots per vertex used by the solver.
constexpr int kLMWorkSlots = 13;

// Per-vertex arrays (all float4, indexed vbase + i).
struct LMArrays {
  const float4* x0;     // initial normalized position (xyz), w = sum of edge multiplicities
  float4* u;            // displacement from x0 (solution), in/out
  const int2* adj;      // (first entry, count) of the vertex's adjacency list
  const unsigned* nbr;  // entry: local neighbour index (low 24 bits) | multiplicity << 24
  float4* work;         // kLMWorkSlots float4 per vertex of scratch, slot-major
  const float* wt;      // per adjacency entry: squared edge weight (weighted pairs only)
};

// General kernel (vectors in global memory): pairs plist[0..npairs) of pairs_dev.
int lm_solve_launch(const LMPairDesc* pairs_dev, const int* plist, int npairs, const LMArrays& arrays,
                    const LMOptionsDev& opt, LMResultDev* results_dev, long long total_vertices, int block_threads,
                    hipStream_t s);

// Multi-workgroup kernel (k_lm_solve_team): team t of wg_dev's entries
// (team, rank, size, -) solves pair plist[t] on `size` co-resident
// workgroups (cooperative launch, nwg <= lm_team_capacity).  cnt: one
// zeroed counter per team at a 128-byte stride (32 unsigned); part: 2 *
// lm_team_max() * 8 doubles per team; abort_flag: one zeroed int, set when a
// barrier wait exceeded its bound.
// lm_team_launch / lm_rot_team_launch return this (> 0, not an error) when
// the runtime refuses the cooperative launch; nothing was launched and the
// caller runs the one-workgroup kernel instead.
constexpr int kCoopRefused = 1;
int lm_team_max();
int lm_team_part_doubles();   // per-team FP64 partial storage (all team kernels)
int lm_team_capacity(int block_threads);
// Held from a multi-workgroup (cooperative) launch until its stream is
// synchronized: two cooperative grids on different streams must not wait on
// each other's residency.
std::mutex& team_launch_mutex();
int lm_team_launch(const LMPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg,
                   const LMArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                   long long total_vertices, int block_threads, unsigned* cnt, double* part, int* abort_flag,
                   hipStream_t s);

// Register/LDS-resident kernel (mo_lm_lds.hip) for pairs of at most 8192
// vertices whose adjacency multiplicities are 1 or 2, in size buckets.
int lm_lds_buckets();
int lm_lds_bucket_for(int nv);                       // -1: too large
int lm_lds_shape(int bucket, int* threads, int* vertices_per_thread);
int lm_lds_set_table(const int* max_nv, const int* shape, int n);   // n = 0: built-in order
size_t lm_lds_bytes(int nv, long long entries);      // dynamic LDS for one pair
// next: one device int per launch, zero before it (the launch's work counter)
int lm_lds_launch(int bucket, const LMPairDesc* pairs_dev, const int* plist, int count, int* next,
                  const LMArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                  long long total_vertices, size_t lds_bytes, hipStream_t s);

// Multi-workgroup register/LDS kernel (k_lm_solve_lds_team) for pairs
// above one workgroup's capacity with multiplicities <= 2: workgroup r of a
// team owns a contiguous chunk of the pair's vertices.  Team shape b (0..2)
// holds 4096, 3072 or 2048 vertices per workgroup; lm_lds_chunk_entries
// writes, per pair, the largest chunk's adjacency entry count for each
// shape (3 ints per pair); lm_lds_team_bytes gives the dynamic LDS of a
// chunk; wg_dev entries are (team, rank, team size, chunk); cnt / part /
// abort_flag as for lm_team_launch.  Returns kCoopRefused like
// lm_team_launch.
int lm_lds_team_shapes();
// Team shapes in preference order (MO_LM_TEAM_VPT overrides the default
// largest-chunk-first order).
const std::vector<int>& lm_lds_team_order();
int lm_lds_team_max();   // largest team of the register/LDS team kernel
int lm_lds_team_shape(int b, int* threads, int* vertices_per_thread);
size_t lm_lds_team_bytes(int chunk, long long entries);
int lm_lds_team_capacity(int b, size_t lds_bytes);
int lm_lds_chunk_entries(const LMPairDesc* pairs_dev, const int2* adj, int np, int* out_dev, hipStream_t s);
int lm_lds_team_launch(int b, const LMPairDesc* pairs_dev, const int* plist, const int4* wg_dev, int nwg,
                       const LMArrays& arrays, const LMOptionsDev& opt, LMResultDev* results_dev,
                       long long total_vertices, size_t lds_bytes, unsigned* cnt, double* part, int* abort_flag,
                       hipStream_t s);

}  // namespace mo
