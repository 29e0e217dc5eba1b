// rigid_rot_deform — drop-in for the reference CLI (src/app/rigid_rot_deform.cc:12-63)
// on libmeshode's C ABI.
//
//   rigid_rot_deform source.obj reference.obj output.obj [GRID_RESOLUTION=64]
//                    [MESH_RESOLUTION=5000] [lambda=1] [symmetry=0]
//
// Kept from the reference: argc < 5 prints the usage line and exits 0
// (:13-18); MESH_RESOLUTION is parsed and unused (:28-29); symmetry mirrors
// the reference mesh in its raw coordinates (:36-41); the size lines
// (:44-47); the reference normalized, the source moved into its frame and
// the distance grid built (:49-54, on the GPU); Deformer::DeformWithRot
// (deformer.cc:94-171, on the GPU: positions and per-vertex angle-axis
// rotations) with its three cost lines (deformer.cc:152-170); "Deformed"
// (:59) and the source written in the reference's original frame (:61,
// mesh.cc:46-64).  Ceres' per-iteration progress table is replaced by one
// summary line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <vector>

#include "../../include/meshode.h"
#include "obj_io.h"

namespace {

int die(const char* what) {
  std::cerr << "rigid_rot_deform: " << what << ": " << mo_last_error() << std::endl;
  return 1;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::cerr << "rigid_rot_deform: " #x ": " << hipGetErrorString(e_) << std::endl;    \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

template <class T>
int upload(T** d, const std::vector<T>& h) {
  HIPCHK(hipMalloc((void**)d, std::max<size_t>(h.size(), 4) * sizeof(T)));
  if (!h.empty()) HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

const char* termination_name(int t) {
  return t == 0 ? "CONVERGENCE" : t == 1 ? "NO_CONVERGENCE" : "FAILURE";
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    printf("./rigid_rot_deform source.obj reference.obj output.obj "
           "[GRID_RESOLUTION=64] [MESH_RESOLUTION=5000]"
           "[lambda=1] [symmetry=0]\n");
    return 0;
  }
  mo_cli::Mesh src, ref;
  if (!mo_cli::read_obj(argv[1], &src)) { std::cerr << "cannot read " << argv[1] << std::endl; return 1; }
  if (!mo_cli::read_obj(argv[2], &ref)) { std::cerr << "cannot read " << argv[2] << std::endl; return 1; }
  int GRID_RESOLUTION = 64, MESH_RESOLUTION = 5000;
  if (argc > 4) sscanf(argv[4], "%d", &GRID_RESOLUTION);
  if (argc > 5) sscanf(argv[5], "%d", &MESH_RESOLUTION);
  (void)MESH_RESOLUTION;
  double lambda = 1;
  if (argc > 6) sscanf(argv[6], "%lf", &lambda);
  int symmetry = 0;
  if (argc > 7) sscanf(argv[7], "%d", &symmetry);
  if (symmetry) mo_cli::reflection_symmetrize(&ref);

  std::cout << "Source:\t\t" << "Num vertices: " << src.nv() << "\tNum faces: " << src.nf() << std::endl;
  std::cout << "Reference:\t" << "Num vertices: " << ref.nv() << "\tNum faces: " << ref.nf() << std::endl
            << std::endl;

  double *dref = nullptr, *dV = nullptr;
  int32_t *drf = nullptr, *dF = nullptr;
  if (upload(&dref, ref.V) || upload(&drf, ref.F) || upload(&dV, src.V) || upload(&dF, src.F)) return 1;
  int32_t pid = -1;
  if (mo_template_create_f64(dref, (int64_t)ref.nv(), drf, (int64_t)ref.nf(), 0, GRID_RESOLUTION, nullptr, &pid) !=
      MO_OK)
    return die("reference grid");
  if (mo_apply_transform_f64(dV, (int64_t)src.nv(), pid, 0, nullptr) != MO_OK) return die("transform");

  // DeformWithRot's defaults (inner PCG tolerance 1e-6) with the one problem
  // spread over a team of 64-thread workgroups, 512 vertices each (a
  // 5000-vertex solve: 276 ms against 747 ms on one workgroup)
  mo_lm_options opt;
  mo_lm_default_options(&opt);
  opt.pcg_relative_tolerance = 1e-6f;
  opt.block_threads = 64;
  opt.team_vertices = 512;
  mo_lm_stats st;
  if (mo_deform_rot(pid, dV, (int64_t)src.nv(), dF, (int64_t)src.nf(), lambda, &opt, nullptr, &st, nullptr) != MO_OK)
    return die("DeformWithRot");
  std::cout << "LM: " << st.iterations << " iterations (" << st.successful_steps << " successful), "
            << st.pcg_iterations << " PCG iterations, " << termination_name(st.termination) << ", solver "
            << mo_cli::solver_path_name(st.solver_path) << std::endl;
  mo_cli::warn_if_degraded(st.solver_path);
  std::cout << "Vertices cost: " << st.vertex_cost << std::endl;
  std::cout << "Rigidity cost: " << st.rigidity_cost << std::endl;
  std::cout << "Final cost: " << st.vertex_cost + st.rigidity_cost << std::endl;
  std::cout << "Deformed" << std::endl;

  if (mo_apply_transform_f64(dV, (int64_t)src.nv(), pid, 1, nullptr) != MO_OK) return die("transform");
  std::vector<double> V(src.V.size());
  HIPCHK(hipMemcpy(V.data(), dV, V.size() * sizeof(double), hipMemcpyDeviceToHost));
  mo_cli::write_obj(argv[3], V, src.F);
  mo_template_destroy(pid);
  for (void* p : {(void*)dref, (void*)dV, (void*)drf, (void*)dF}) (void)hipFree(p);
  return 0;
}
