// C ABI of the CAD chain: preprocessing handles (host), Delaunay helpers
// (host), FP64 template transforms and LinearEstimation (device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mo_cad.h"
#include "mo_common.h"
#include "mo_delaunay.h"
#include "mo_linear.h"

#define MO_EXPORT extern "C" __attribute__((visibility("default")))

namespace mo {
namespace {

struct CadHandle {
  cad::Mesh clean;
  cad::Subdivision sub;
};

std::mutex g_cad_mu;
std::vector<std::unique_ptr<CadHandle>> g_cad;

CadHandle* get_cad(int64_t h) {
  std::lock_guard<std::mutex> lk(g_cad_mu);
  if (h < 0 || h >= (int64_t)g_cad.size() || !g_cad[h]) return nullptr;
  return g_cad[h].get();
}

__global__ void k_xform64(double* V, long long n3, double scale, double t0, double t1, double t2, int inverse) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const int k = (int)(i % 3);
  const double t = k == 0 ? t0 : (k == 1 ? t1 : t2);
  V[i] = inverse ? V[i] * scale + t : (V[i] - t) / scale;
}

}  // namespace
}  // namespace mo

using namespace mo;

MO_EXPORT int mo_cad_prepare(const double* V, int64_t nv, const int32_t* F, int64_t nf, double len_subdivide,
                             double step_neighbors, double step_graph, int32_t threads, int64_t* handle) {
  MO_CHECK(V && F && handle && nv > 0 && nf >= 0 && nv < (1LL << 30) && nf < (1LL << 30), MO_ERR_ARG,
           "mo_cad_prepare: bad arguments");
  MO_CHECK(len_subdivide > 0 && step_neighbors > 0 && step_graph > 0, MO_ERR_ARG,
           "mo_cad_prepare: lengths must be positive");
  for (int64_t i = 0; i < 3 * nf; ++i)
    MO_CHECK(F[i] >= 0 && F[i] < nv, MO_ERR_RANGE, "mo_cad_prepare: face index %d out of range", F[i]);
  auto h = std::make_unique<CadHandle>();
  h->clean.V.assign(V, V + 3 * nv);
  h->clean.F.assign(F, F + 3 * nf);
  cad::remove_degenerated(&h->clean);
  cad::merge_duplex(&h->clean);
  if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  try {
    cad::subdivide(h->clean, len_subdivide, &h->sub);
    cad::geometry_neighbors(&h->sub, step_neighbors, threads);
    cad::representative_graph(&h->sub, step_graph);
  } catch (const std::exception& e) {
    return fail(MO_ERR_NOMEM, "mo_cad_prepare: %s", e.what());
  }
  std::lock_guard<std::mutex> lk(g_cad_mu);
  g_cad.push_back(std::move(h));
  *handle = (int64_t)g_cad.size() - 1;
  return MO_OK;
}

MO_EXPORT int mo_cad_sizes(int64_t handle, int64_t* sizes7) {
  CadHandle* h = get_cad(handle);
  MO_CHECK(h && sizes7, MO_ERR_ARG, "mo_cad_sizes: unknown handle %lld", (long long)handle);
  sizes7[0] = (int64_t)h->sub.mesh.nv();
  sizes7[1] = (int64_t)h->sub.mesh.nf();
  sizes7[2] = (int64_t)h->sub.neighbors.size();
  sizes7[3] = (int64_t)h->sub.graph_v.size() / 3;
  sizes7[4] = (int64_t)h->sub.graph_e.size();
  sizes7[5] = (int64_t)h->clean.nv();
  sizes7[6] = (int64_t)h->clean.nf();
  return MO_OK;
}

MO_EXPORT int mo_cad_export(int64_t handle, double* V, int32_t* F, int32_t* E, int32_t* ref, double* GV, int32_t* GE,
                            double* clean_V, int32_t* clean_F) {
  CadHandle* h = get_cad(handle);
  MO_CHECK(h, MO_ERR_ARG, "mo_cad_export: unknown handle %lld", (long long)handle);
  const cad::Subdivision& s = h->sub;
  if (V) std::copy(s.mesh.V.begin(), s.mesh.V.end(), V);
  if (F) std::copy(s.mesh.F.begin(), s.mesh.F.end(), F);
  if (E)
    for (size_t i = 0; i < s.neighbors.size(); ++i) {
      E[2 * i] = s.neighbors[i].first;
      E[2 * i + 1] = s.neighbors[i].second;
    }
  if (ref) std::copy(s.ref.begin(), s.ref.end(), ref);
  if (GV) std::copy(s.graph_v.begin(), s.graph_v.end(), GV);
  if (GE)
    for (size_t i = 0; i < s.graph_e.size(); ++i) {
      GE[2 * i] = s.graph_e[i].first;
      GE[2 * i + 1] = s.graph_e[i].second;
    }
  if (clean_V) std::copy(h->clean.V.begin(), h->clean.V.end(), clean_V);
  if (clean_F) std::copy(h->clean.F.begin(), h->clean.F.end(), clean_F);
  return MO_OK;
}

MO_EXPORT int mo_cad_destroy(int64_t handle) {
  std::lock_guard<std::mutex> lk(g_cad_mu);
  MO_CHECK(handle >= 0 && handle < (int64_t)g_cad.size() && g_cad[handle], MO_ERR_ARG,
           "mo_cad_destroy: unknown handle %lld", (long long)handle);
  g_cad[handle].reset();
  return MO_OK;
}

MO_EXPORT int mo_delaunay2(const double* xy, int64_t n, int32_t* tris, int64_t cap, int64_t* ntri) {
  MO_CHECK(xy && ntri && n >= 0 && n < (1LL << 30), MO_ERR_ARG, "mo_delaunay2: bad arguments");
  std::vector<int> t;
  geom::delaunay2(xy, (int)n, &t);
  *ntri = (int64_t)t.size() / 3;
  MO_CHECK((int64_t)t.size() <= 3 * cap, MO_ERR_RANGE, "mo_delaunay2: %lld triangles exceed capacity %lld",
           (long long)(t.size() / 3), (long long)cap);
  if (tris) std::copy(t.begin(), t.end(), tris);
  return MO_OK;
}

MO_EXPORT int mo_delaunay3(const double* xyz, int64_t n, int32_t* simplices, int64_t cap, int64_t* count,
                           int32_t* dim) {
  MO_CHECK(xyz && count && dim && n >= 0 && n < (1LL << 30), MO_ERR_ARG, "mo_delaunay3: bad arguments");
  std::vector<int> s;
  int d = 0;
  geom::delaunay3(xyz, (int)n, &d, &s);
  *dim = d;
  const int k = d >= 1 ? d + 1 : 1;
  *count = (int64_t)s.size() / k;
  MO_CHECK((int64_t)s.size() <= cap * 4, MO_ERR_RANGE, "mo_delaunay3: output exceeds capacity");
  if (simplices) std::copy(s.begin(), s.end(), simplices);
  return MO_OK;
}

MO_EXPORT int mo_apply_transform_f64(double* V, int64_t nv, int32_t param_id, int32_t inverse, void* stream) {
  MO_CHECK(nv >= 0, MO_ERR_ARG, "mo_apply_transform_f64: bad vertex count");
  if (nv == 0) return MO_OK;
  MO_CHECK(V, MO_ERR_ARG, "mo_apply_transform_f64: null array");
  int N = 0;
  double scale = 1, t[3] = {0, 0, 0};
  if (mo_template_info(param_id, &N, &scale, t) != MO_OK) return MO_ERR_ARG;
  const long long n3 = 3 * (long long)nv;
  hipLaunchKernelGGL(k_xform64, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, V, n3, scale,
                     t[0], t[1], t[2], inverse ? 1 : 0);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

MO_EXPORT int mo_linear_estimation(double* V, int64_t nv, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                                   const int32_t* ref, const double* GV, int64_t ngv, double rigidity,
                                   int32_t* iterations, void* stream) {
  MO_CHECK(V && ref && GV && nv > 0 && ngv > 0 && nf >= 0 && ne >= 0 && (F || nf == 0) && (E || ne == 0), MO_ERR_ARG,
           "mo_linear_estimation: bad arguments");
  int it = 0;
  const int rc = linear_estimation(V, nv, F, nf, E, ne, ref, GV, ngv, rigidity, 100000, 1e-13, &it,
                                   (hipStream_t)stream);
  if (iterations) *iterations = it;
  return rc;
}

MO_EXPORT int mo_linear_estimation_rot(double* V, int64_t nv, const int32_t* F, int64_t nf, const double* TV,
                                       double rigidity, int32_t* iterations, void* stream) {
  MO_CHECK(V && TV && nv > 0 && nf >= 0 && (F || nf == 0), MO_ERR_ARG, "mo_linear_estimation_rot: bad arguments");
  int it = 0;
  const int rc = linear_estimation_rot(V, nv, F, nf, TV, rigidity, 100000, 1e-13, &it, (hipStream_t)stream);
  if (iterations) *iterations = it;
  return rc;
}
