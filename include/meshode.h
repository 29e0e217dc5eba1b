/*
 * meshode.h — C ABI of libmeshode.so, the MI355X (gfx950) implementation of
 * MeshODE's deformation hot path.
 *
 * Conventions
 *   - Every function returns an int status: 0 on success, a negative MO_ERR_*
 *     code on failure; mo_last_error() returns the message of the calling
 *     thread's last failure.
 *   - Bulk arrays are DEVICE pointers (HIP memory of the current device),
 *     row-major and contiguous; sizes are element counts.  `stream` is a
 *     hipStream_t (NULL = default stream).  Calls are stream-ordered; calls
 *     that return host values (ids, scalars, stats) synchronize the stream.
 *   - Vertex arrays are [n,3]; faces [n,3] int32 (0-based); edges [n,2] int32.
 *   - No torch types cross this boundary.
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to the hjwdzh/MeshODE checkout).
 */
#ifndef MESHODE_H_
#define MESHODE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MO_OK 0
#define MO_ERR_ARG (-1)
#define MO_ERR_HIP (-2)
#define MO_ERR_STATE (-3)
#define MO_ERR_RANGE (-4)
#define MO_ERR_NOMEM (-5)
#define MO_ERR_NOCONV (-6)

/* Library version (major * 10000 + minor * 100 + patch). */
int mo_version(void);

/* Message of the calling thread's last failed call ("" if none). */
const char* mo_last_error(void);

/* ------------------------------------------------------------------------
 * Templates: the per-target state of the pyDeform parameter store.
 * ---------------------------------------------------------------------- */

/* InitializeDeformTemplate(tensorV, tensorF, symmetry, grid_resolution)
 * (src/interface/deform_params.h:30-34, deform_params.cc:16-40).
 * Copies the target, normalizes it (Mesh::Normalize, mesh.cc:66-96) and
 * builds the N^3 distance grid (Mesh::ConstructDistanceField,
 * mesh.cc:106-152).  `symmetry` is accepted and ignored, as in the
 * reference (it symmetrizes the still-empty mesh, deform_params.cc:28-29).
 * Writes the new parameter id (ids count up from 0 like the reference's
 * g_params index).  Synchronizes `stream`. */
int mo_template_create_f32(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t symmetry,
                           int32_t grid_resolution, void* stream, int32_t* param_id);

/* Same from FP64 vertices (the CLI path: Mesh::ReadOBJ stores doubles,
 * rigid_deform.cc:49-52). */
int mo_template_create_f64(const double* V, int64_t nv, const int32_t* F, int64_t nf, int32_t symmetry,
                           int32_t grid_resolution, void* stream, int32_t* param_id);

/* Releases a template's device memory (the reference never frees its
 * g_params slots, deform_params.cc:7-14).  Ids are not reused. */
int mo_template_destroy(int32_t param_id);

/* Rebuilds an existing template in place for a new FP64 target: the same
 * normalization and grid build as mo_template_create_f64
 * (deform_params.cc:16-40 repeated on an existing slot), into the
 * template's buffers at its grid resolution, with no device allocation or
 * free for the grid (hipFree synchronizes the whole device).  Stored edge
 * offsets are dropped.  For a serving cache that recycles evicted targets'
 * slots.  The caller must not have work using the template in flight.
 * Synchronizes `stream`. */
int mo_template_rebuild_f64(int32_t param_id, const double* V, int64_t nv, const int32_t* F, int64_t nf,
                            void* stream);

/* Grid resolution and the normalization of a template:
 * DeformParams::scale / trans (deform_params.h:15-16). */
int mo_template_info(int32_t param_id, int32_t* grid_resolution, double* scale, double* trans3);

/* Copies the template's grid to device buffers (any may be NULL):
 * dist64 = UniformGrid voxels (uniformgrid.h:33) as [z][y][x] FP64,
 * dist32 = the same rounded to float, nearest = nearest triangle index
 * (igl::point_mesh_squared_distance's I, discarded by mesh.cc:137-150;
 * lowest index among equal distances).  A template grid is exact up to the
 * band b = 0.2 + sqrt(3)/N + 0.01 and holds b (nearest -1) beyond it: the
 * lookups truncate interpolated values above 0.2 (uniformgrid.cc:80-82) and
 * the field is 1-Lipschitz, so no lookup result, value or gradient, depends
 * on a value beyond b.  mo_grid_build computes the exact field everywhere. */
int mo_template_grid(int32_t param_id, double* dist64, float* dist32, int32_t* nearest, void* stream);

/* ------------------------------------------------------------------------
 * Distance-field build primitive.
 * ---------------------------------------------------------------------- */

/* Mesh::ConstructDistanceField (mesh.cc:106-152) on already-normalized
 * FP64 vertices: dist[(i*N+j)*N+k] = distance from (k/N, j/N, i/N) to the
 * nearest triangle.  Outputs may be NULL. */
int mo_grid_build(const double* Vn, int64_t nv, const int32_t* F, int64_t nf, int32_t grid_resolution,
                  double* dist64, float* dist32, int32_t* nearest, void* stream);

/* ------------------------------------------------------------------------
 * pyDeform layer functions (src/interface/pydeform.cc:19-37).
 * ---------------------------------------------------------------------- */

/* NormalizeByTemplate / DenormalizeByTemplate (normalize.h:6-10,
 * normalize.cc:5-45): in place on FP32 vertices. */
int mo_normalize_by_template(float* V, int64_t nv, int32_t param_id, void* stream);
int mo_denormalize_by_template(float* V, int64_t nv, int32_t param_id, void* stream);

/* DistanceFieldLoss_forward (distance_layer.h:8-10, distance_layer.cc:8-37):
 * out[i] = d(V_i)^2, out [nv]. */
int mo_distance_loss_forward(const float* V, int64_t nv, int32_t param_id, float* out, void* stream);

/* DistanceFieldLoss_backward (distance_layer.h:12-14, distance_layer.cc:39-80):
 * out[i] = d(V_i) * grad d(V_i), out [nv,3]. */
int mo_distance_loss_backward(const float* V, int64_t nv, int32_t param_id, float* out, void* stream);

/* Edge losses (rigid and graph, forward and backward): a template checks an
 * index buffer (F or E) on the host the first time it sees it (MO_ERR_RANGE
 * for an index outside [0, nv)) and keeps the buffer's incidence lists by
 * its device address; later calls with that address neither copy nor wait
 * on the stream.  A buffer whose contents change must come at a new address
 * (or be stored again). */

/* StoreRigidityInformation (rigid_layer.h:9-12, rigid_layer.cc:11-46). */
int mo_rigidity_store(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id, void* stream);

/* RigidEdgeLoss_forward (rigid_layer.h:14-17, rigid_layer.cc:48-89): out [3nf,3]. */
int mo_rigid_loss_forward(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id, float* out,
                          void* stream);

/* RigidEdgeLoss_backward (rigid_layer.h:19-22, rigid_layer.cc:91-133): out [nv,3]. */
int mo_rigid_loss_backward(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id, float* out,
                           void* stream);

/* StoreGraphInformation (graph_layer.h:9-12, graph_layer.cc:11-45). */
int mo_graph_store(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id, void* stream);

/* GraphEdgeLoss_forward (graph_layer.h:14-17, graph_layer.cc:47-91): out [ne,3]. */
int mo_graph_loss_forward(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id, float* out,
                          void* stream);

/* GraphEdgeLoss_backward (graph_layer.h:19-22, graph_layer.cc:93-135): out [nv,3]. */
int mo_graph_loss_backward(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id, float* out,
                           void* stream);

/* ------------------------------------------------------------------------
 * Levenberg-Marquardt deformation (Deformer, src/lib/deformer.h:11-30).
 * ---------------------------------------------------------------------- */

typedef struct mo_lm_options {
  int32_t max_num_iterations;            /* 100 (deformer.cc:56) */
  int32_t max_consecutive_invalid_steps; /* 5 */
  double function_tolerance;             /* 1e-6 */
  double gradient_tolerance;             /* 1e-10 */
  double parameter_tolerance;            /* 1e-8 */
  double min_relative_decrease;          /* 1e-3 */
  double initial_trust_region_radius;    /* 1e4 */
  double max_trust_region_radius;        /* 1e16 */
  double min_trust_region_radius;        /* 1e-32 */
  double min_lm_diagonal;                /* 1e-6 */
  double max_lm_diagonal;                /* 1e32 */
  int32_t pcg_max_iterations;            /* inner PCG cap per LM step (1000) */
  float pcg_relative_tolerance;          /* ||r|| <= tol ||r0|| ends PCG (1e-5) */
  int32_t block_threads;                 /* threads per pair of the global-memory kernel (1024) */
  int32_t timing;                        /* nonzero: record phase times (mo_last_timings); 2: also
                                            count the grid build's work (mo_last_grid_stats; costs a
                                            few percent of the grid phase) */
  int32_t kernel_variant;                /* 0: register/LDS kernel when a pair fits, else the
                                            global-memory kernels (a team of workgroups per pair
                                            above team_vertices); 1: one-workgroup global-memory
                                            kernel only; 2: multi-workgroup kernel for every pair */
  int32_t max_successful_steps;          /* > 0: stop after that many accepted steps (0) */
  int32_t team_vertices;                 /* pairs on the multi-workgroup kernel: vertices per
                                            workgroup of a pair's team (0: 4 * block_threads);
                                            mo_deform_rot(_batch): > 0 puts each problem on a
                                            team of that many vertices per workgroup (0: one
                                            workgroup per problem) */
} mo_lm_options;

typedef struct mo_lm_stats {
  double initial_cost;
  double final_cost;     /* "Final cost" line, deformer.cc:90 */
  double vertex_cost;    /* "Vertices cost" line, deformer.cc:81 */
  double rigidity_cost;  /* "Rigidity cost" line, deformer.cc:87 */
  int32_t iterations;
  int32_t successful_steps;
  int32_t unsuccessful_steps;
  int32_t termination;   /* 0 convergence, 1 iteration cap, 2 failure */
  int64_t pcg_iterations;
  int64_t clock_start;   /* device wall clock (ticks) when the pair's solve began */
  int64_t clock_end;     /* ... and ended; see mo_last_timings for the rate */
  int32_t solver_path;   /* MO_PATH_*: which kernel solved the pair */
  int32_t workgroups;    /* workgroups that solved it (1, or the team's size) */
} mo_lm_stats;

/* mo_lm_stats.solver_path.  MO_PATH_COOP_REFUSED marks a pair meant for a
 * multi-workgroup team that ran on one workgroup because the runtime refused
 * the cooperative launch (e.g. while another process held the GPU): same
 * algorithm, different reduction order (results differ in low bits) and a
 * longer solve. */
#define MO_PATH_LDS 0            /* register/LDS kernel, one workgroup */
#define MO_PATH_ONE_WORKGROUP 1  /* global-memory kernel, one workgroup */
#define MO_PATH_TEAM 2           /* multi-workgroup (cooperative) kernel */
#define MO_PATH_COOP_REFUSED 3   /* one workgroup after a refused cooperative launch */
#define MO_PATH_LDS_TEAM 4       /* register/LDS kernel on a team of workgroups (cooperative) */

/* Fills the defaults listed above. */
void mo_lm_default_options(mo_lm_options* opt);

/* Launch shapes of the register/LDS LM kernel (no counterpart in the
 * reference, whose Ceres solve is single-threaded, deformer.cc:55-59: the
 * tuning interface of this library's batched solver, for the autotuner in
 * meshode_amd/autotune.py).  Shape i runs one pair on threads[i] threads
 * holding vertices_per_thread[i] vertices each, so it takes pairs of at most
 * threads * vertices_per_thread vertices (and multiplicities 1 and 2).
 * mo_lm_set_shape_table installs a size-class table: a pair of nv vertices
 * uses shape[r] for the first r (max_vertices ascending) with
 * nv <= max_vertices[r]; a larger pair, or every pair when n = 0 (the
 * default), takes the first fitting shape of the built-in order
 * (MO_LM_SHAPES="TxVPT,..." overrides that order).  Each row's shape must
 * hold max_vertices[r] vertices.  The table is process-wide and read under a
 * lock by every later mo_deform_batch / mo_rigid_deform_batch call. */
int32_t mo_lm_shape_count(void);
int mo_lm_shape_info(int32_t index, int32_t* threads, int32_t* vertices_per_thread);
int mo_lm_set_shape_table(const int32_t* max_vertices, const int32_t* shape, int32_t n);

/* Deformer::Deform (deformer.cc:18-92; kind 0, idx = faces [n,3]) or
 * Deformer::DeformGraph (deformer.cc:370-442; kind 1, idx = edges [n,2]) or
 * the same with AdaptiveEdgeLoss residuals (edgeloss.h:35-61; kind 2, idx =
 * edges [n,2], weight lambda * 2e-2 / (|v| + 1e-8) per edge) for a batch of
 * pairs in one persistent launch.  Pair p deforms V[p] ([nv[p],3]
 * FP64, already in the template's normalized frame, updated in place)
 * against the grid of template param_ids[p] with rigidity lambda[p].
 * param_ids, V, nv, idx, nidx, kind, lambda and stats are HOST arrays of
 * length npairs (V and idx hold device pointers).  Synchronizes `stream`. */
int mo_deform_batch(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                    const int32_t* const* idx, const int64_t* nidx, const int32_t* kind, const double* lambda,
                    const mo_lm_options* opt, mo_lm_stats* stats, void* stream);

/* mo_deform_batch without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight serving
 * loops, code that must not call hipMalloc/hipFree).  Same arguments,
 * kernels and results as mo_deform_batch; the solver's scratch (per-vertex
 * state, adjacency, team counters and partials, pair descriptors, results)
 * lives in one library-owned device array of mo_deform_batch_noalloc_max_bytes()
 * bytes (64 MiB) per GPU, kept until the process ends.  A batch that needs
 * more returns MO_ERR_RANGE before anything is launched (about 1.8 MB per
 * 5000-vertex, 30k-half-edge pair: ~35 such pairs).  options.timing must be 0.
 * Calls on one GPU share its array: within a process they are serialized
 * (each returns after its device work has finished); the array keeps the
 * last call's contents, and whatever a call needs zeroed it clears itself. */
int64_t mo_deform_batch_noalloc_max_bytes(void);
int mo_deform_batch_noalloc(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                            const int32_t* const* idx, const int64_t* nidx, const int32_t* kind,
                            const double* lambda, const mo_lm_options* opt, mo_lm_stats* stats, void* stream);

/* The whole rigid_deform pipeline (rigid_deform.cc:49-61) for a batch of
 * pairs, device-resident end to end: normalize target p
 * (Mesh::Normalize), build its grid_resolution[p]^3 distance field, bring
 * source p into the target frame (Mesh::ApplyTransform), solve
 * Deformer::Deform with lambda[p], and write the deformed source in the
 * target's original frame (Mesh::WriteOBJ's v * scale + pos, mesh.cc:46-64)
 * to out[p] ([src_nv[p],3] FP64, may alias src[p]).  All per-pair
 * arguments are HOST arrays of length npairs holding device pointers /
 * sizes.  stats (host, optional) receives the per-pair solver summary.
 * Synchronizes `stream`. */
int mo_rigid_deform_batch(int32_t npairs, const double* const* src, const int64_t* src_nv,
                          const int32_t* const* src_faces, const int64_t* src_nf, const double* const* tgt,
                          const int64_t* tgt_nv, const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                          const int32_t* grid_resolution, const double* lambda, const mo_lm_options* opt,
                          double* const* out, mo_lm_stats* stats, void* stream);

/* The same pipeline as a captured HIP graph (no counterpart in the
 * reference, which runs one ceres::Solve per process, rigid_deform.cc:12-63:
 * the launch-chain replay for callers that solve the same batch shape again
 * and again).  mo_rigid_deform_batch_graph runs the batch once, exactly as
 * mo_rigid_deform_batch does (out and stats are that run's), then captures
 * the whole chain -- target normalization, grid builds, source transform,
 * topology, LM, write-back -- into a graph with its own device scratch and
 * pinned staging (allocated here, once), and returns it in *graph.
 * mo_batch_graph_launch replays it on `stream` and waits for it: every
 * kernel runs again on the current contents of the same src / tgt buffers
 * (positions may change between replays; faces, sizes, resolutions and
 * lambdas are the captured ones) and writes the same out buffers, and stats
 * (optional) receive that replay's solver statistics.  A replay checks the
 * values that steered the captured chain (adjacency entry counts, error
 * flags) and returns MO_ERR_STATE when they differ.  out[p] must not alias
 * src[p]; options.timing must be 0; a batch with a pair that needs a
 * cooperative multi-workgroup launch (MO_PATH_TEAM / MO_PATH_LDS_TEAM) is
 * not captured (MO_ERR_ARG after the eager run, whose results are valid).
 * A graph belongs to the device current at capture; replays of one graph
 * must not overlap.  mo_batch_graph_destroy frees it. */
typedef struct mo_batch_graph mo_batch_graph;
int mo_rigid_deform_batch_graph(int32_t npairs, const double* const* src, const int64_t* src_nv,
                                const int32_t* const* src_faces, const int64_t* src_nf, const double* const* tgt,
                                const int64_t* tgt_nv, const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                                const int32_t* grid_resolution, const double* lambda, const mo_lm_options* opt,
                                double* const* out, mo_lm_stats* stats, void* stream, mo_batch_graph** graph);
int mo_batch_graph_launch(mo_batch_graph* graph, mo_lm_stats* stats, void* stream);
int mo_batch_graph_destroy(mo_batch_graph* graph);

/* mo_rigid_deform_batch without the final wait (no counterpart in the
 * reference, which runs one ceres::Solve per process, rigid_deform.cc:12-63:
 * for callers that keep several batches in flight on several streams, so
 * that one batch's grid builds fill the CUs that the previous batch's last
 * solves leave idle).  The call enqueues the whole pipeline on `stream`
 * and returns once the solver launches are queued (it waits only for the
 * sources' topology, whose sizes steer those launches); *handle then owns
 * the batch.  after (optional): a handle not yet waited for, on another
 * stream; this batch's grid builds start only once that batch's solver
 * launches are about to start, so the solvers are placed on the CUs first
 * and the grid builds take the CUs they leave.  mo_batch_wait(handle,
 * stats) waits for the batch, fills stats (optional; host, npairs entries)
 * exactly as mo_rigid_deform_batch would, reports a bad target face index
 * as that call would, and frees the handle (call it once per handle, from
 * any thread).  out[p] is written when the batch completes and the
 * caller's buffers must stay alive until then.  options.timing must be 0.
 * A batch with a pair that needs a cooperative multi-workgroup launch
 * (MO_PATH_TEAM / MO_PATH_LDS_TEAM) completes inside the call. */
typedef struct mo_batch_handle mo_batch_handle;
int mo_rigid_deform_batch_async(int32_t npairs, const double* const* src, const int64_t* src_nv,
                                const int32_t* const* src_faces, const int64_t* src_nf, const double* const* tgt,
                                const int64_t* tgt_nv, const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                                const int32_t* grid_resolution, const double* lambda, const mo_lm_options* opt,
                                double* const* out, void* stream, mo_batch_handle* after,
                                mo_batch_handle** handle);
int mo_batch_wait(mo_batch_handle* handle, mo_lm_stats* stats);

/* Deformer::DeformSubdivision with a TerminateWhenSuccessCallback
 * (deformer.cc:173-263, callback.h:9-39, as cad_deform.cc:87-114 runs it):
 * distance residuals on every vertex of the subdivided mesh plus
 * AdaptiveEdgeLoss residuals on the geometry-neighbour pairs E [ne,2] and
 * the face half-edges of F [nf,3], rest offsets from V_rest; the solve is
 * repeated from the current positions until two consecutive final costs
 * differ by less than 1e-6 (at most max_solves times), each solve stopping
 * at the callback's second successful step.  V [nv,3] FP64 (device; start
 * positions in, result out; may equal V_rest only before the first call).
 * At the first successful step of a solve the callback receives the
 * positions the solve started from (Ceres hands the user's parameter blocks
 * back only when a solve ends), as a device pointer valid during the call.
 * stats (host, optional) receives the last solve's summary, *solves (host,
 * optional) the number of solves.  Synchronizes `stream`. */
typedef void (*mo_flow_callback)(const double* V, int64_t nv, void* user);
int mo_deform_subdivision(int32_t param_id, const double* V_rest, double* V, int64_t nv, const int32_t* F,
                          int64_t nf, const int32_t* E, int64_t ne, double lambda, const mo_lm_options* opt,
                          int32_t max_solves, mo_flow_callback callback, void* user, mo_lm_stats* stats,
                          int32_t* solves, void* stream);

/* Phase times of the calling thread's last mo_deform_batch /
 * mo_rigid_deform_batch call made with opt->timing set, from HIP events on
 * the call's stream: out[0] distance-grid builds (ms), out[1] topology and
 * source setup (ms), out[2] the LM solver kernel (ms), out[3] whole call
 * (ms), out[4] device wall clock (ticks) at the start of the call, out[5]
 * wall clock rate (Hz).  No counterpart in the reference (its only timing
 * is Ceres' progress output, deformer.cc:58). */
int mo_last_timings(double* out6);

/* Work counters of the calling thread's last mo_rigid_deform_batch call made
 * with opt->timing = 2 (no counterpart in the reference): out4 = point-
 * triangle distance evaluations executed by the grid build (64 lanes per
 * wavefront evaluation), the evaluations some lane needed, face bound checks
 * executed (64 lanes each) and face tiles staged, summed over the batch. */
int mo_last_grid_stats(int64_t* out4);

/* ---------------------------------------------------------------------------
 * CAD chain (cad_deform.cc:21-120, LoadCadMesh mesh_tensor.cc:102-178).
 * The preprocessing is host work, as in the reference; arrays of the
 * mo_cad_* and mo_delaunay* functions are HOST arrays. */

/* Mesh::RemoveDegenerated + Mesh::MergeDuplex (mesh.cc:320-333, 181-232),
 * Subdivision::Subdivide(len_subdivide) (subdivision.cc:28-250),
 * ComputeGeometryNeighbors(step_neighbors) (:252-341) and
 * ComputeRepresentativeGraph(step_graph) (:343-399) of the CAD mesh V
 * [nv,3] FP64 / F [nf,3] (raw coordinates).  The reference uses 2e-2,
 * 1.5e-2, 1e-2 (cad_deform.cc:46-48).  threads <= 0 picks the host's core
 * count (at most 16).  *handle identifies the result until mo_cad_destroy. */
int mo_cad_prepare(const double* V, int64_t nv, const int32_t* F, int64_t nf, double len_subdivide,
                   double step_neighbors, double step_graph, int32_t threads, int64_t* handle);

/* sizes7: subdivided vertices, subdivided faces, geometry-neighbour pairs,
 * representative vertices, representative edges, cleaned CAD vertices,
 * cleaned CAD faces. */
int mo_cad_sizes(int64_t handle, int64_t* sizes7);

/* Copies the results (any pointer may be NULL): subdivided mesh V [n,3]
 * FP64 and F [m,3]; E [e,2] the geometry-neighbour pairs
 * (Subdivision::Neighbors); ref [n] the vertex -> representative map
 * (Vertex2Graph); GV [g,3] FP64 and GE [h,2] the representative graph
 * (GraphV, GraphE); the cleaned CAD mesh.  Pairs are (i < j), sorted, as the
 * reference's std::set iterates them. */
int mo_cad_export(int64_t handle, double* V, int32_t* F, int32_t* E, int32_t* ref, double* GV, int32_t* GE,
                  double* clean_V, int32_t* clean_F);

int mo_cad_destroy(int64_t handle);

/* Delaunay2D (delaunay.cc:50-74): triangles (counter-clockwise) of the 2D
 * points xy [n,2]; *ntri receives the count, tris [cap,3]. */
int mo_delaunay2(const double* xy, int64_t n, int32_t* tris, int64_t cap, int64_t* ntri);

/* Delaunay3D (delaunay.cc:76-101): simplices of the 3D points xyz [n,3];
 * *dim = 3 (tetrahedra), 2 (coplanar: triangles), 1 (collinear: segments),
 * 0/-1 (nothing); *count simplices of dim+1 indices, simplices [cap,4]. */
int mo_delaunay3(const double* xyz, int64_t n, int32_t* simplices, int64_t cap, int64_t* count, int32_t* dim);

/* Mesh::ApplyTransform (mesh.cc:98-105; Subdivision::ApplyTransform,
 * subdivision.cc:19-26) with template param_id's normalization: V [nv,3]
 * FP64 device array in place, v = (v - trans) / scale; inverse != 0 maps
 * back, v = v * scale + trans (Mesh::WriteOBJ, mesh.cc:46-64). */
int mo_apply_transform_f64(double* V, int64_t nv, int32_t param_id, int32_t inverse, void* stream);

/* LinearEstimation (linear.cc:11-112; Subdivision::LinearSolve,
 * subdivision.cc:462-471; pyDeform SolveLinear with_rot = 0,
 * linear_layer.cc:5-87): V [nv,3] FP64 rest positions in, solution out;
 * F [nf,3], E [ne,2] (geometry-neighbour pairs), ref [nv], GV [ngv,3] FP64
 * (deformed representative vertices); all device arrays.  The SPD system is
 * solved by FP64 conjugate gradients on the device to a relative residual
 * of 1e-13 (MO_ERR_NOCONV otherwise).  *iterations (host, optional)
 * receives the CG iteration count.  Synchronizes `stream`. */
int mo_linear_estimation(double* V, int64_t nv, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                         const int32_t* ref, const double* GV, int64_t ngv, double rigidity, int32_t* iterations,
                         void* stream);

/* Deformer::DeformWithRot (deformer.cc:94-171; rigid_rot_deform.cc:55-56):
 * positions and per-vertex angle-axis rotations, DistanceLoss per vertex and
 * EdgeLossWithRot per face half-edge (a, b) = (F[f][j], F[f][j+1]),
 *   lambda (rotate(w_a, p_a - p_b) - (V_a - V_b)),  w_a - w_b
 * (edgeloss.h:64-98; half-edges with a == b are skipped), by the LM loop of
 * mo_deform_batch (Ceres' defaults, max 100 iterations; opt = null takes
 * the defaults with pcg_relative_tolerance 1e-6, tighter than Deform's
 * 1e-5: with 1e-5 the inexact steps can settle in a neighbouring minimum,
 * 1e-3 above the exact-step solve's energy).  V [nv,3] FP64
 * device array in the template's normalized frame: the rest shape in, the
 * solution out.  rot [nv,3] FP64 device array or null: the start rotations
 * (zeros when null, as the reference) in, the final ones out.  F [nf,3]
 * device.  opt.block_threads above 256 is clamped to 256.  Synchronizes
 * `stream`. */
int mo_deform_rot(int32_t param_id, double* V, int64_t nv, const int32_t* F, int64_t nf, double lambda,
                  const mo_lm_options* opt, double* rot, mo_lm_stats* stats, void* stream);

/* mo_deform_rot for a batch of independent problems in one launch (one
 * workgroup each): pair p deforms V[p] ([nv[p],3] FP64 device, normalized
 * frame of template param_ids[p]) with faces F[p] ([nf[p],3]) and weight
 * lambda[p]; rot (nullable array, entries nullable) as in mo_deform_rot;
 * stats[p] per pair.  Each pair's result does not depend on the others. */
int mo_deform_rot_batch(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                        const int32_t* const* F, const int64_t* nf, const double* lambda, const mo_lm_options* opt,
                        double* const* rot, mo_lm_stats* stats, void* stream);

/* LinearEstimationWithRot (linear.cc:114-230; linear.h:14-15; pyDeform
 * SolveLinear with_rot = 1, linear_layer.cc:76-78): V [nv,3] FP64 rest
 * positions in, solution out; F [nf,3] int32; TV [nv,3] FP64 target
 * positions; all device arrays.  Per-vertex frames s_i R_i (scaled polar
 * factor of the neighbourhood covariance) and the right-hand side are built
 * on the device, then the SPD system (unit diagonal plus the face-edge
 * Laplacian with weights 2e-2 rigidity / |e|) is solved by FP64 CG to a
 * relative residual of 1e-13 (MO_ERR_NOCONV otherwise).  Vertices joined
 * by a zero-length face edge (the reference's weight is infinite there and
 * its right-hand side NaN) are solved as one unknown, the limit as the
 * edge length goes to 0.  MO_ERR_ARG for a face index out of range.
 * Synchronizes `stream`. */
int mo_linear_estimation_rot(double* V, int64_t nv, const int32_t* F, int64_t nf, const double* TV,
                             double rigidity, int32_t* iterations, void* stream);

/* Nearest source point of each target point (ReverseLossLayer's
 * cKDTree(src).query(tar, k = 1), reverse_loss_layer.py:18-19): src [ns,3],
 * tar [nt,3] FP32 device arrays; index [nt] int32 and, if non-null,
 * dist2 [nt] FP64 squared distances (device); tar and index may be null
 * when nt = 0.  Distances are FP64 from the
 * FP32 coordinates, (dx*dx + dy*dy) + dz*dz; ties go to the lowest index. */
int mo_nearest_f32(const float* src, int64_t ns, const float* tar, int64_t nt, int32_t* index, double* dist2,
                   void* stream);

/* ---- NeuralODE flow: one fixed-grid RK4 step of ODEFunc -----------------
 * Replaces the torch evaluation of odeint(ODEFunc, y0, [t0, t1],
 * method="rk4") (neuralode_fast.py:7-33, 42-61; torchdiffeq's 3/8-rule
 * step) and its autograd backward.  ODEFunc is 4 -> 50 -> 50 -> 50 -> 3,
 * LeakyReLU 0.01, input cat(y, t) - 0.5.  params: the module's 5503
 * parameters, flat in its parameter order (W0 [50,4], b0, W1 [50,50], b1,
 * W2 [50,50], b2, W3 [3,50], b3), FP32 device.  tau4 (host): the network's
 * time input at the four stages; sign: +1, or -1 for a decreasing grid
 * (integrated as t -> -t with f -> -f(-t, y), so tau = -stage time); dt:
 * the step in the (possibly negated) grid.  y0, y1 [n,3] FP32 device.
 *
 * mo_ode_rk4_forward: y1 = y0 + (k1 + 3 (k2 + k3) + k4) dt / 8.  zsave
 * [n,9] (optional, device) receives the stage inputs z2, z3, z4 for the
 * backward.
 * mo_ode_rk4_backward: from gy1 = dL/dy1 writes gy0 = dL/dy0 (optional)
 * and gparams [5503] = dL/dparams (overwritten).  work: device workspace
 * of mo_ode_rk4_workspace_bytes(n) bytes (8-byte aligned), or null (then
 * stream-ordered allocation).  Sums over points are formed per 64-point
 * chunk (FP32) and added in a fixed order in FP64: repeated calls give
 * identical results. */
int64_t mo_ode_rk4_workspace_bytes(int64_t n);
int mo_ode_rk4_forward(const float* y0, int64_t n, const float* params, const float* tau4, float sign, float dt,
                       float* y1, float* zsave, void* stream);
int mo_ode_rk4_backward(const float* y0, int64_t n, const float* params, const float* tau4, float sign, float dt,
                        const float* zsave, const float* gy1, float* gy0, float* gparams, void* work,
                        int64_t work_bytes, void* stream);

/* ---- config-3 velocity field: one RK4 step on FP32 MFMA -------------------
 * The field v(y) = MLP(y - 0.5), 3 -> 256 -> 256 -> 256 -> 3 with softplus
 * (torch's threshold 20), the widened ODEFunc of neuralode_fast.py:7-33,
 * and the 3/8-rule step of odeint(method="rk4") (neuralode_fast.py:55-60):
 * y1 = y0 + h (k1 + 3 (k2 + k3) + k4) / 8.  Replaces the torch evaluation of
 * a forward step (meshode_amd/neuralode.py's rk4_step).
 * mo_flow256_pack: the module's tensors (W0 [256,3], b0 [256], W1 [256,256],
 * b1, W2 [256,256], b2, W3 [3,256], b3 [3]; FP32 device, row-major) into
 * `packed` (mo_flow256_packed_floats() floats, device, 16-byte aligned).
 * mo_flow256_step: y0, y1 [n,3] FP32 device (distinct buffers). */
int64_t mo_flow256_packed_floats(void);
int mo_flow256_pack(const float* W0, const float* b0, const float* W1, const float* b1, const float* W2,
                    const float* b2, const float* W3, const float* b3, float* packed, void* stream);
int mo_flow256_step(const float* y0, int64_t n, const float* packed, float h, float* y1, void* stream);
/* The adjoint of a step (the reference differentiates odeint by autograd).
 * mo_flow256_step_save: mo_flow256_step that also writes the stage inputs
 * x = z - 0.5 to Xs [4][n][3] and the activations H1, H2, H3 to
 * Hs [3][4][n][256] (layer, stage, point).  mo_flow256_step_backward: from
 * those and gy1 = dL/dy1 writes gy0 = dL/dy0, the pre-activation gradients
 * dA1, dA2, dA3 to dAs [3][4][n][256] and dL/dk_s to Gs [4][n][3]; the
 * parameter gradients are then dW0 = dA1^T X, dW1 = dA2^T H1,
 * dW2 = dA3^T H2, dW3 = G^T H3 and the column sums of dA1, dA2, dA3, G for
 * the biases (rows = the 4n stage points).  Hs and dAs 16-byte aligned. */
int mo_flow256_step_save(const float* y0, int64_t n, const float* packed, float h, float* y1, float* Hs, float* Xs,
                         void* stream);
int mo_flow256_step_backward(int64_t n, const float* packed, float h, const float* Hs, const float* gy1, float* gy0,
                             float* dAs, float* Gs, void* stream);
/* The small parameter gradients of a step from mo_flow256_step_backward's
 * outputs: out [mo_flow256_small_grad_floats() = 2307] = dW0 [256][3], db0,
 * db1, db2 [256 each], dW3 [3][256], db3 [3]; work: 256 x 2307 floats
 * (device).  Sums in a fixed order (FP64 across workgroups). */
int64_t mo_flow256_small_grad_floats(void);
int mo_flow256_small_grads(int64_t n, const float* dAs, const float* Hs, const float* Xs, const float* Gs, float* out,
                           float* work, void* stream);
/* dW1 = dA2^T H1 and dW2 = dA3^T H2 of a step (sums over its 4n stage
 * points; split-K on MFMA, workgroup partials added in a fixed order):
 * out [2][256][256]; work: mo_flow256_dw_work_floats() floats (device). */
int64_t mo_flow256_dw_work_floats(void);
int mo_flow256_dw_grads(int64_t n, const float* dAs, const float* Hs, float* out, float* work, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MESHODE_H_ */
