// rigid_deform — drop-in for the reference CLI (src/app/rigid_deform.cc:12-63)
// running the deformation on the GPU through libmeshode's C ABI.
//
//   rigid_deform source.obj reference.obj output.obj [GRID_RESOLUTION=64]
//                [MESH_RESOLUTION=5000] [lambda=1] [symmetry=0]
//
// Kept from the reference: argc < 5 prints the usage line and exits 0
// (rigid_deform.cc:13-18); MESH_RESOLUTION is parsed and unused (:36-37);
// symmetry mirrors the reference mesh in its raw coordinates before
// normalization (:25-31, mesh.cc:234-246); the size lines (:44-47), the
// three cost lines of Deformer::Deform (deformer.cc:81-91) and "Deformed"
// (:59); the output mesh is written in the reference's original frame with
// Mesh::WriteOBJ's format (mesh.cc:46-64).  Ceres' per-iteration progress
// table (deformer.cc:58) is replaced by one summary line of this solver.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/meshode.h"
#include "obj_io.h"

namespace {

using mo_cli::Mesh;
using mo_cli::read_obj;
using mo_cli::reflection_symmetrize;
using mo_cli::write_obj;

int die(const char* what) {
  std::cerr << "rigid_deform: " << what << ": " << mo_last_error() << std::endl;
  return 1;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::cerr << "rigid_deform: " #x ": " << hipGetErrorString(e_) << std::endl;        \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    printf("./rigid_deform source.obj reference.obj output.obj "
           "[GRID_RESOLUTION=64] [MESH_RESOLUTION=5000] "
           "[lambda=1] [symmetry=0]\n");
    return 0;
  }
  int GRID_RESOLUTION = 64, MESH_RESOLUTION = 5000, symmetry = 0;
  Mesh src, ref;
  if (!read_obj(argv[1], &src)) { std::cerr << "cannot read " << argv[1] << std::endl; return 1; }
  if (!read_obj(argv[2], &ref)) { std::cerr << "cannot read " << argv[2] << std::endl; return 1; }
  if (argc > 7) sscanf(argv[7], "%d", &symmetry);
  if (symmetry) reflection_symmetrize(&ref);
  if (argc > 4) sscanf(argv[4], "%d", &GRID_RESOLUTION);
  if (argc > 5) sscanf(argv[5], "%d", &MESH_RESOLUTION);
  (void)MESH_RESOLUTION;
  double lambda = 1;
  if (argc > 6) sscanf(argv[6], "%lf", &lambda);

  std::cout << "Source:\t\t" << "Num vertices: " << src.nv() << "\tNum faces: " << src.nf() << std::endl;
  std::cout << "Reference:\t" << "Num vertices: " << ref.nv() << "\tNum faces: " << ref.nf() << std::endl
            << std::endl;

  double *dsrc = nullptr, *dref = nullptr, *dout = nullptr;
  int *dsf = nullptr, *drf = nullptr;
  HIPCHK(hipMalloc(&dsrc, std::max<size_t>(src.V.size(), 3) * sizeof(double)));
  HIPCHK(hipMalloc(&dout, std::max<size_t>(src.V.size(), 3) * sizeof(double)));
  HIPCHK(hipMalloc(&dref, std::max<size_t>(ref.V.size(), 3) * sizeof(double)));
  HIPCHK(hipMalloc(&dsf, std::max<size_t>(src.F.size(), 3) * sizeof(int)));
  HIPCHK(hipMalloc(&drf, std::max<size_t>(ref.F.size(), 3) * sizeof(int)));
  HIPCHK(hipMemcpy(dsrc, src.V.data(), src.V.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dref, ref.V.data(), ref.V.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dsf, src.F.data(), src.F.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(drf, ref.F.data(), ref.F.size() * sizeof(int), hipMemcpyHostToDevice));

  mo_lm_options opt;
  mo_lm_default_options(&opt);
  mo_lm_stats st;
  const double* srcs[1] = {dsrc};
  const double* refs[1] = {dref};
  const int32_t* sfs[1] = {dsf};
  const int32_t* rfs[1] = {drf};
  double* outs[1] = {dout};
  const int64_t snv = (int64_t)src.nv(), snf = (int64_t)src.nf(), rnv = (int64_t)ref.nv(), rnf = (int64_t)ref.nf();
  const int32_t N = GRID_RESOLUTION;
  if (mo_rigid_deform_batch(1, srcs, &snv, sfs, &snf, refs, &rnv, rfs, &rnf, &N, &lambda, &opt, outs, &st,
                            nullptr) != MO_OK)
    return die("deformation");

  std::cout << "LM: " << st.iterations << " iterations (" << st.successful_steps << " successful), "
            << st.pcg_iterations << " PCG iterations, "
            << (st.termination == 0 ? "CONVERGENCE" : st.termination == 1 ? "NO_CONVERGENCE" : "FAILURE")
            << ", solver " << mo_cli::solver_path_name(st.solver_path) << std::endl;
  mo_cli::warn_if_degraded(st.solver_path);
  std::cout << "Vertices cost: " << st.vertex_cost << std::endl;
  std::cout << "Rigidity cost: " << st.rigidity_cost << std::endl;
  std::cout << "Final cost: " << st.vertex_cost + st.rigidity_cost << std::endl;
  std::cout << "Deformed" << std::endl;

  std::vector<double> out(src.V.size());
  HIPCHK(hipMemcpy(out.data(), dout, out.size() * sizeof(double), hipMemcpyDeviceToHost));
  write_obj(argv[3], out, src.F);
  (void)hipFree(dsrc);
  (void)hipFree(dref);
  (void)hipFree(dout);
  (void)hipFree(dsf);
  (void)hipFree(drf);
  return 0;
}
