// cad_deform — drop-in for the reference CLI (src/app/cad_deform.cc:21-120)
// on libmeshode's C ABI.
//
//   cad_deform cad.obj reference.obj output.obj [GRID_RESOLUTION=64]
//              [MESH_RESOLUTION=5000] [lambda=1] [symmetry=0] [flow_output]
//
// Kept from the reference: argc < 5 prints the usage line and exits 0
// (:22-27); the reference mesh is mirrored in raw coordinates when symmetry
// is set (:34-41); the CAD mesh is cleaned, subdivided (2e-2), given
// geometry neighbours (1.5e-2) and a representative graph (1e-2) (:43-51,
// host code); the size lines print the cleaned CAD mesh (:60-63); the
// reference is normalized and its distance grid built (:65-70, on the GPU);
// the representative graph is deformed (Deformer::DeformGraph,
// deformer.cc:370-442, on the GPU) and its motion propagated to the
// subdivided mesh by LinearEstimation (Subdivision::LinearSolve,
// subdivision.cc:462-471, on the GPU); "Deformed" (:116) and the subdivided
// mesh written in the reference's original frame (:118, mesh.cc:46-64).
// With flow_output (:76-114) the subdivided mesh itself is deformed
// (Deformer::DeformSubdivision, AdaptiveEdgeLoss, repeated solves that stop
// at the callback's second successful step) and the recorded positions are
// written as one polyline per vertex; no LinearSolve in that mode.
// Ceres' per-iteration progress table is replaced by one summary line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/meshode.h"
#include "obj_io.h"

namespace {

int die(const char* what) {
  std::cerr << "cad_deform: " << what << ": " << mo_last_error() << std::endl;
  return 1;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::cerr << "cad_deform: " #x ": " << hipGetErrorString(e_) << std::endl;          \
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
    printf("./cad_deform cad.obj reference.obj output.obj "
           "[GRID_RESOLUTION=64] [MESH_RESOLUTION=5000] "
           "[lambda=1] [symmetry=0] [flow_output=filename]\n");
    return 0;
  }
  mo_cli::Mesh ref, cad;
  if (!mo_cli::read_obj(argv[1], &cad)) { std::cerr << "cannot read " << argv[1] << std::endl; return 1; }
  if (!mo_cli::read_obj(argv[2], &ref)) { std::cerr << "cannot read " << argv[2] << std::endl; return 1; }
  int symmetry = 0;
  if (argc > 7) sscanf(argv[7], "%d", &symmetry);
  if (symmetry) mo_cli::reflection_symmetrize(&ref);

  int64_t h = -1;
  if (mo_cad_prepare(cad.V.data(), (int64_t)cad.nv(), cad.F.data(), (int64_t)cad.nf(), 2e-2, 1.5e-2, 1e-2, 0, &h) !=
      MO_OK)
    return die("CAD preprocessing");
  int64_t n[7];
  if (mo_cad_sizes(h, n) != MO_OK) return die("CAD preprocessing");
  std::vector<double> V(n[0] * 3), GV(n[3] * 3);
  std::vector<int32_t> F(n[1] * 3), E(n[2] * 2), R(n[0]), GE(n[4] * 2);
  if (mo_cad_export(h, V.data(), F.data(), E.data(), R.data(), GV.data(), GE.data(), nullptr, nullptr) != MO_OK)
    return die("CAD preprocessing");
  mo_cad_destroy(h);

  int GRID_RESOLUTION = 64, MESH_RESOLUTION = 5000;
  if (argc > 4) sscanf(argv[4], "%d", &GRID_RESOLUTION);
  if (argc > 5) sscanf(argv[5], "%d", &MESH_RESOLUTION);
  (void)MESH_RESOLUTION;
  double lambda = 1;
  if (argc > 6) sscanf(argv[6], "%lf", &lambda);

  std::cout << "Source:\t\t" << "Num vertices: " << n[5] << "\tNum faces: " << n[6] << std::endl;
  std::cout << "Reference:\t" << "Num vertices: " << ref.nv() << "\tNum faces: " << ref.nf() << std::endl
            << std::endl;

  double *dref = nullptr, *dV = nullptr, *dGV = nullptr;
  int32_t *drf = nullptr, *dF = nullptr, *dE = nullptr, *dR = nullptr, *dGE = nullptr;
  if (upload(&dref, ref.V) || upload(&drf, ref.F) || upload(&dV, V) || upload(&dF, F) || upload(&dE, E) ||
      upload(&dR, R) || upload(&dGV, GV) || upload(&dGE, GE))
    return 1;
  int32_t pid = -1;
  if (mo_template_create_f64(dref, (int64_t)ref.nv(), drf, (int64_t)ref.nf(), 0, GRID_RESOLUTION, nullptr, &pid) !=
      MO_OK)
    return die("reference grid");
  if (mo_apply_transform_f64(dV, n[0], pid, 0, nullptr) != MO_OK ||
      mo_apply_transform_f64(dGV, n[3], pid, 0, nullptr) != MO_OK)
    return die("transform");

  mo_lm_options opt;
  mo_lm_default_options(&opt);
  if (argc > 8) {
    // Deformer::DeformSubdivision on the subdivided mesh, recording the
    // flow at each solve's first successful step (cad_deform.cc:87-114)
    struct Flow {
      std::vector<std::vector<double>> snaps;
    } flow;
    auto record = [](const double* Vd, int64_t nv, void* user) {
      std::vector<double> h((size_t)nv * 3);
      if (hipMemcpy(h.data(), Vd, h.size() * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess)
        static_cast<Flow*>(user)->snaps.push_back(std::move(h));
    };
    double* dRest = nullptr;
    HIPCHK(hipMalloc(&dRest, V.size() * sizeof(double)));
    HIPCHK(hipMemcpy(dRest, dV, V.size() * sizeof(double), hipMemcpyDeviceToDevice));
    mo_lm_stats st;
    int32_t nsolve = 0;
    if (mo_deform_subdivision(pid, dRest, dV, n[0], dF, n[1], dE, n[2], lambda, &opt, 100000, record, &flow, &st,
                              &nsolve, nullptr) != MO_OK)
      return die("DeformSubdivision");
    std::vector<double> last(V.size());
    HIPCHK(hipMemcpy(last.data(), dV, last.size() * sizeof(double), hipMemcpyDeviceToHost));
    flow.snaps.push_back(last);
    std::cout << "LM: " << nsolve << " solves, last " << st.iterations << " iterations (" << st.successful_steps
              << " successful), " << termination_name(st.termination) << std::endl;
    std::cout << "Vertices cost: " << st.vertex_cost << std::endl;
    std::cout << "Rigidity cost: " << st.rigidity_cost << std::endl;
    std::cout << "Final cost: " << st.vertex_cost + st.rigidity_cost << std::endl;
    int32_t gN = 0;
    double scale = 1, pos[3] = {0, 0, 0};
    if (mo_template_info(pid, &gN, &scale, pos) != MO_OK) return die("template");
    std::ofstream os(argv[8]);
    long long top = 1;
    const size_t ns = flow.snaps.size();
    for (int64_t i = 0; i < n[0]; ++i) {
      for (size_t j = 0; j < ns; ++j) {
        const double* v = &flow.snaps[j][3 * i];
        os << "v " << v[0] * scale + pos[0] << " " << v[1] * scale + pos[1] << " " << v[2] * scale + pos[2] << "\n";
      }
      for (size_t j = 0; j + 1 < ns; ++j) os << "l " << top + (long long)j << " " << top + (long long)j + 1 << "\n";
      top += (long long)ns;
    }
    os.close();
    (void)hipFree(dRest);
  } else {
    // Deformer::DeformGraph on the representative graph
    mo_lm_stats st;
    const int32_t kind = 1;
    double* gvs[1] = {dGV};
    const int32_t* ges[1] = {dGE};
    if (mo_deform_batch(1, &pid, gvs, &n[3], ges, &n[4], &kind, &lambda, &opt, &st, nullptr) != MO_OK)
      return die("DeformGraph");
    std::cout << "LM: " << st.iterations << " iterations (" << st.successful_steps << " successful), "
              << st.pcg_iterations << " PCG iterations, " << termination_name(st.termination) << ", solver "
              << mo_cli::solver_path_name(st.solver_path) << std::endl;
    mo_cli::warn_if_degraded(st.solver_path);
    std::cout << "Vertices cost: " << st.vertex_cost << std::endl;
    std::cout << "Rigidity cost: " << st.rigidity_cost << std::endl;
    std::cout << "Final cost: " << st.vertex_cost + st.rigidity_cost << std::endl;
    // Subdivision::LinearSolve (default rigidity 2, linear.h:8-14)
    int32_t cg_it = 0;
    if (mo_linear_estimation(dV, n[0], dF, n[1], dE, n[2], dR, dGV, n[3], 2.0, &cg_it, nullptr) != MO_OK)
      return die("LinearSolve");
  }
  std::cout << "Deformed" << std::endl;

  if (mo_apply_transform_f64(dV, n[0], pid, 1, nullptr) != MO_OK) return die("transform");
  HIPCHK(hipMemcpy(V.data(), dV, V.size() * sizeof(double), hipMemcpyDeviceToHost));
  mo_cli::write_obj(argv[3], V, F);
  mo_template_destroy(pid);
  for (void* p : {(void*)dref, (void*)dV, (void*)dGV, (void*)drf, (void*)dF, (void*)dE, (void*)dR, (void*)dGE})
    (void)hipFree(p);
  return 0;
}
