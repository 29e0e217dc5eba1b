// C ABI of libmeshode.so (declared in include/meshode.h): the template
// registry (the pyDeform parameter store, deform_params.cc:7-14), the
// pyDeform layer entry points, and the batched LM deformation pipeline.
#include "../../include/meshode.h"

#include "mo_common.h"
#include "mo_grid.h"
#include "mo_layers.h"
#include "mo_lm.h"
#include "mo_lm_rot.h"
#include "mo_topo.h"

#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#define MO_EXPORT extern "C" __attribute__((visibility("default")))

namespace mo {

static thread_local std::string g_last_error;
static thread_local double g_timings[6] = {0, 0, 0, 0, 0, 0};
static thread_local long long g_grid_stats[kGridStatCount] = {0, 0, 0, 0};

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// ---------------------------------------------------------------------------
// Batch graphs: a batch's launch chain (normalization, grid build, topology,
// LM, write-back) is run twice by mo_rigid_deform_batch_graph -- once
// eagerly in record mode, then under hipStreamBeginCapture in capture mode --
// and the captured graph is replayed by mo_batch_graph_launch.  The chain's
// host code reaches the stream through these helpers where the modes differ:
//   Arena::reserve  record: sizes summed; capture: carved from the graph's
//                   device pool (no stream allocation inside the graph)
//   chain_h2d       capture: the bytes are staged in the graph's pinned
//                   buffer first (a replay copies the same bytes)
//   chain_read      device values the host reads to steer the chain (entry
//                   counts, error flags).  eager: copy + synchronize; record:
//                   the same, values kept in order; capture: the recorded
//                   values are returned and the device copy goes to the
//                   pinned buffer, compared with the record after each replay
// The LM results are copied to the pinned buffer at the end of the graph.
struct ChainCtx {
  enum Mode { kRecord, kCapture };
  Mode mode = kRecord;
  std::vector<std::vector<char>> reads;   // record: steering values, in order
  size_t pool_bytes = 0;                  // record: device scratch of the chain
  size_t pinned_bytes = 0;                // record: staged copies + reads + results
  bool team_launch = false;               // record: a cooperative team launch was made
  size_t next_read = 0;                   // capture
  char* pool = nullptr;
  size_t pool_cap = 0, pool_used = 0;
  char* pinned = nullptr;
  size_t pinned_cap = 0, pinned_used = 0;
  std::vector<std::pair<size_t, size_t>> checks;   // (pinned offset, index into reads)
  size_t res_off = 0;
  int res_np = 0;
  std::vector<int> res_path, res_wgs;
};

thread_local ChainCtx* g_chain = nullptr;

// Set while mo_rigid_deform_batch_async enqueues a batch: the LM results
// and the error flag are copied into `pinned` (page-locked, so the copies
// stay asynchronous) instead of being read back, and the call returns with
// the work in flight.  deferred stays false when the batch needed a
// cooperative team launch: that path completes inside the call.
struct AsyncCtx {
  char* pinned = nullptr;
  size_t err_off = 0;
  bool deferred = false;
  int np = 0;
  std::vector<int> path, wgs;
  hipEvent_t lm_ready = nullptr;    // recorded just before the batch's solver launches
  hipEvent_t after = nullptr;       // the grid builds wait for this event (another batch's lm_ready)
  ~AsyncCtx() {
    if (lm_ready) (void)hipEventDestroy(lm_ready);
  }
};

thread_local AsyncCtx* g_async = nullptr;

// page-locked blocks, reused across calls (freeing page-locked memory
// synchronizes the device, which would stall the batches in flight)
std::mutex g_pinned_mu;
std::vector<std::pair<size_t, char*>> g_pinned_free;

int pinned_get(size_t bytes, char** p, size_t* cap) {
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    for (size_t k = 0; k < g_pinned_free.size(); ++k)
      if (g_pinned_free[k].first >= bytes) {
        *cap = g_pinned_free[k].first;
        *p = g_pinned_free[k].second;
        g_pinned_free.erase(g_pinned_free.begin() + (long)k);
        return MO_OK;
      }
  }
  const size_t c = std::max<size_t>((bytes + 65535) & ~(size_t)65535, 65536);
  MO_HIP(hipHostMalloc((void**)p, c, hipHostMallocDefault));
  *cap = c;
  return MO_OK;
}

void pinned_put(char* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pinned_mu);
  g_pinned_free.emplace_back(cap, p);
}

// Page-locked staging of one batch call's host <-> device copies, so that
// they stay asynchronous (a copy from pageable memory returns only once it
// has run, i.e. once the stream has drained up to it).  The blocks go back
// to the pool when the call ends -- for an asynchronous batch, when it is
// waited for -- after the stream has passed every copy that reads them.
struct HostStage {
  hipStream_t stream = nullptr;
  std::vector<std::pair<char*, size_t>> blocks;
  size_t used = 0;                  // bytes taken from the last block
  hipEvent_t ev = nullptr;          // read-back fence
  bool drained = false;             // set when the stream is known to be past every copy
  explicit HostStage(hipStream_t s) : stream(s) {}
  ~HostStage() {
    if (!drained && !blocks.empty()) (void)hipStreamSynchronize(stream);
    if (ev) (void)hipEventDestroy(ev);
    for (auto& b : blocks) pinned_put(b.first, b.second);
  }
  int take(size_t bytes, char** p) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (blocks.empty() || used + bytes > blocks.back().second) {
      char* b = nullptr;
      size_t cap = 0;
      int rc = pinned_get(std::max<size_t>(bytes, 1 << 20), &b, &cap);
      if (rc) return rc;
      blocks.emplace_back(b, cap);
      used = 0;
    }
    *p = blocks.back().first + used;
    used += bytes;
    return MO_OK;
  }
};

thread_local HostStage* g_stage = nullptr;

struct StageScope {
  explicit StageScope(HostStage* st) { g_stage = st; }
  ~StageScope() { g_stage = nullptr; }
};

// Second streams for work of one call that may overlap its main stream's
// (a batch's source topology beside its grid builds): taken from a pool per
// device, returned when the call ends, kept for the process's lifetime.
std::mutex g_aux_mu;
std::vector<std::pair<int, hipStream_t>> g_aux_free;

struct AuxStream {
  int device = -1;
  hipStream_t s = nullptr;
  int acquire() {
    MO_HIP(hipGetDevice(&device));
    {
      std::lock_guard<std::mutex> lk(g_aux_mu);
      for (size_t k = 0; k < g_aux_free.size(); ++k)
        if (g_aux_free[k].first == device) {
          s = g_aux_free[k].second;
          g_aux_free.erase(g_aux_free.begin() + (long)k);
          return MO_OK;
        }
    }
    MO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return MO_OK;
  }
  ~AuxStream() {
    if (!s) return;
    // nothing of this call may still run on it when the call's scratch and
    // staging are released (a no-op once its sizes have been read)
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(g_aux_mu);
    g_aux_free.emplace_back(device, s);
  }
};

inline size_t al256c(size_t b) { return (b + 255) & ~(size_t)255; }

static int pinned_take(ChainCtx* c, size_t bytes, size_t* off) {
  MO_CHECK(c->pinned_used + al256c(bytes) <= c->pinned_cap, MO_ERR_STATE,
           "batch graph capture: pinned staging exceeds the recorded size");
  *off = c->pinned_used;
  c->pinned_used += al256c(bytes);
  return MO_OK;
}

int chain_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return MO_OK;
  ChainCtx* c = g_chain;
  if (c && c->mode == ChainCtx::kCapture) {
    size_t off = 0;
    int rc = pinned_take(c, bytes, &off);
    if (rc) return rc;
    memcpy(c->pinned + off, src, bytes);
    src = c->pinned + off;
  } else {
    if (c) c->pinned_bytes += al256c(bytes);
    if (HostStage* st = g_stage) {
      char* p = nullptr;
      int rc = st->take(bytes, &p);
      if (rc) return rc;
      memcpy(p, src, bytes);
      src = p;
    }
  }
  MO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  return MO_OK;
}

struct ChainRead {
  void* host;
  const void* dev;
  size_t bytes;
};

// Reads n device values into host memory.  between(), when given, is
// called after the reads are enqueued and before the host waits for them:
// it enqueues work that the reads do not depend on (so that it runs while
// the host waits and decides).
static int chain_read(const ChainRead* r, int n, hipStream_t s, const std::function<int()>& between = {}) {
  ChainCtx* c = g_chain;
  if (c && c->mode == ChainCtx::kCapture) {
    for (int k = 0; k < n; ++k) {
      MO_CHECK(c->next_read < c->reads.size() && c->reads[c->next_read].size() == r[k].bytes, MO_ERR_STATE,
               "batch graph capture diverged from the recorded run");
      memcpy(r[k].host, c->reads[c->next_read].data(), r[k].bytes);
      size_t off = 0;
      int rc = pinned_take(c, r[k].bytes, &off);
      if (rc) return rc;
      MO_HIP(hipMemcpyAsync(c->pinned + off, r[k].dev, r[k].bytes, hipMemcpyDeviceToHost, s));
      c->checks.emplace_back(off, c->next_read);
      ++c->next_read;
    }
    return between ? between() : MO_OK;
  }
  int rc = MO_OK;
  if (HostStage* st = g_stage) {
    std::vector<char*> tmp(n);
    for (int k = 0; k < n; ++k) {
      if ((rc = st->take(r[k].bytes, &tmp[k]))) return rc;
      MO_HIP(hipMemcpyAsync(tmp[k], r[k].dev, r[k].bytes, hipMemcpyDeviceToHost, s));
    }
    if (!st->ev) MO_HIP(hipEventCreateWithFlags(&st->ev, hipEventDisableTiming));
    MO_HIP(hipEventRecord(st->ev, s));
    if (between && (rc = between())) return rc;
    MO_HIP(hipEventSynchronize(st->ev));
    for (int k = 0; k < n; ++k) memcpy(r[k].host, tmp[k], r[k].bytes);
  } else {
    for (int k = 0; k < n; ++k) MO_HIP(hipMemcpyAsync(r[k].host, r[k].dev, r[k].bytes, hipMemcpyDeviceToHost, s));
    if (between && (rc = between())) return rc;
    MO_HIP(hipStreamSynchronize(s));
  }
  if (c)
    for (int k = 0; k < n; ++k) {
      const char* h = (const char*)r[k].host;
      c->reads.emplace_back(h, h + r[k].bytes);
      c->pinned_bytes += al256c(r[k].bytes);
    }
  return MO_OK;
}

namespace {

// ---------------------------------------------------------------------------
// Device memory helpers.

struct DevPtr {
  void* p = nullptr;
  DevPtr() = default;
  DevPtr(const DevPtr&) = delete;
  DevPtr& operator=(const DevPtr&) = delete;
  ~DevPtr() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  int alloc(size_t bytes) {
    reset();
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(MO_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return MO_OK;
  }
  template <class T> T* as() const { return (T*)p; }
};

// Stream-ordered scratch arena carved into aligned slices: one
// hipMallocAsync (reserve), or a caller-provided region that is neither
// allocated nor freed here (attach).
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  hipStream_t stream = nullptr;
  bool owned = false;
  ~Arena() {
    if (base && owned) (void)hipFreeAsync(base, stream);
  }
  int reserve(size_t bytes, hipStream_t s) {
    stream = s;
    size = bytes + 256;
    if (ChainCtx* c = g_chain) {
      if (c->mode == ChainCtx::kCapture) {
        MO_CHECK(c->pool_used + al256c(size) <= c->pool_cap, MO_ERR_STATE,
                 "batch graph capture: scratch exceeds the recorded size");
        base = c->pool + c->pool_used;
        c->pool_used += al256c(size);
        owned = false;
        return MO_OK;
      }
      c->pool_bytes += al256c(size);
    }
    hipError_t e = hipMallocAsync((void**)&base, size, s);
    if (e != hipSuccess) {
      base = nullptr;
      return fail(MO_ERR_NOMEM, "hipMallocAsync(%zu) failed: %s", size, hipGetErrorString(e));
    }
    owned = true;
    return MO_OK;
  }
  int attach(char* region, size_t cap, size_t bytes, hipStream_t s) {
    MO_CHECK(bytes + 256 <= cap, MO_ERR_RANGE,
             "batch needs %zu bytes of solver scratch; the no-allocation entry point holds at most %zu "
             "(split the batch, or call mo_deform_batch)",
             bytes + 256, cap);
    stream = s;
    base = region;
    size = cap;
    owned = false;
    return MO_OK;
  }
  size_t end() const { return (used + 255) & ~(size_t)255; }
  template <class T> T* take(size_t count) {
    size_t off = (used + 255) & ~(size_t)255;
    used = off + count * sizeof(T);
    return (T*)(base + off);
  }
};

inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// ---------------------------------------------------------------------------
// Template registry.

struct Template {
  int N = 0;
  int device = 0;
  double scale = 1.0;
  double trans[3] = {0, 0, 0};
  long long cells = 0;
  DevPtr grid64, grid32, nearest;
  // edge offsets shared by the Rigid / Graph layers (DeformParams::edge_offset,
  // deform_params.h:18-20; the last Store call wins)
  DevPtr edge_off;
  long long edge_count = 0;
  int edge_kind = -1;
};

std::mutex g_mu;
// Never destroyed: templates left at process exit are not freed by a static
// destructor (hipFree during static destruction may run after the HIP
// runtime or a profiling tool has shut down); the process's exit releases
// their device memory.  mo_template_destroy frees a template explicitly.
std::vector<std::unique_ptr<Template>>& g_templates = *new std::vector<std::unique_ptr<Template>>();

Template* get_template(int id) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (id < 0 || id >= (int)g_templates.size() || !g_templates[id]) return nullptr;
  return g_templates[id].get();
}

int register_template(std::unique_ptr<Template> t) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_templates.push_back(std::move(t));
  return (int)g_templates.size() - 1;
}

hipStream_t as_stream(void* s) { return (hipStream_t)s; }

__global__ void k_zero_u32(unsigned* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

int zero_fill(void* p, size_t bytes, hipStream_t s) {
  long long n = (long long)((bytes + 3) / 4);
  if (n <= 0) return MO_OK;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (unsigned*)p, n);
  MO_HIP(hipGetLastError());
  return MO_OK;
}

// Normalize the target (Mesh::Normalize) and build the template's grid
// (Mesh::ConstructDistanceField) into t's existing buffers.
template <class T>
int template_fill(Template* t, const T* V, int64_t nv, const int32_t* F, int64_t nf, hipStream_t s) {
  MO_CHECK(V && F, MO_ERR_ARG, "template vertices/faces must be non-null device pointers");
  MO_CHECK(nv > 0 && nf > 0, MO_ERR_ARG, "template needs vertices and faces (nv=%lld nf=%lld)", (long long)nv,
           (long long)nf);
  int rc;
  GridSpec gs{nullptr, nv, F, nf, t->N, t->grid64.as<double>(), t->grid32.as<float>(), t->nearest.as<int>(),
              lookup_band2(t->N)};
  const size_t gws = grid_batch_workspace_bytes(&gs, 1);
  Arena ar;
  if ((rc = ar.reserve(al(nv * 3 * sizeof(double)) + al(8 * sizeof(double)) + al(xform_workspace_bytes(1)) + al(gws) +
                           256,
                       s)))
    return rc;
  double* Vn = ar.take<double>(nv * 3);
  double* sp = ar.take<double>(8);
  int* err = (int*)(sp + 4);
  void* xws = ar.take<char>(xform_workspace_bytes(1));
  void* ws = ar.take<char>(gws);
  gs.Vn = Vn;
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  XformJob xj{(const void*)V, sizeof(T) == 8 ? 1 : 0, nv, sp, Vn};
  if ((rc = xform_batch_launch(&xj, 1, true, xws, s))) return rc;
  if ((rc = grid_build_batch_launch(&gs, 1, ws, s, err))) return rc;
  double host[5];
  MO_HIP(hipMemcpyAsync(host, sp, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
  MO_HIP(hipStreamSynchronize(s));
  int herr;
  std::memcpy(&herr, &host[4], sizeof(int));
  MO_CHECK(herr == 0, MO_ERR_RANGE, "template face index out of range [0, %lld)", (long long)nv);
  t->scale = host[0];
  t->trans[0] = host[1];
  t->trans[1] = host[2];
  t->trans[2] = host[3];
  return MO_OK;
}

template <class T>
int template_create(const T* V, int64_t nv, const int32_t* F, int64_t nf, int32_t grid_resolution, void* stream,
                    int32_t* param_id) {
  MO_CHECK(param_id, MO_ERR_ARG, "param_id output pointer is null");
  MO_CHECK(V && F, MO_ERR_ARG, "template vertices/faces must be non-null device pointers");
  MO_CHECK(nv > 0 && nf > 0, MO_ERR_ARG, "template needs vertices and faces (nv=%lld nf=%lld)", (long long)nv,
           (long long)nf);
  MO_CHECK(grid_resolution >= 2 && grid_resolution <= 1024, MO_ERR_ARG, "grid_resolution %d out of range [2, 1024]",
           grid_resolution);
  hipStream_t s = as_stream(stream);
  auto t = std::make_unique<Template>();
  MO_HIP(hipGetDevice(&t->device));
  t->N = grid_resolution;
  t->cells = (long long)grid_resolution * grid_resolution * grid_resolution;
  int rc;
  if ((rc = t->grid64.alloc(t->cells * sizeof(double)))) return rc;
  if ((rc = t->grid32.alloc(t->cells * sizeof(float)))) return rc;
  if ((rc = t->nearest.alloc(t->cells * sizeof(int)))) return rc;
  if ((rc = template_fill<T>(t.get(), V, nv, F, nf, s))) return rc;
  *param_id = register_template(std::move(t));
  return MO_OK;
}

int check_err_flag(int* err_dev, hipStream_t s, const char* what) {
  int h = 0;
  const ChainRead rd{&h, err_dev, sizeof(int)};
  int rc = chain_read(&rd, 1, s);
  if (rc) return rc;
  MO_CHECK(h == 0, MO_ERR_RANGE, "%s: index out of range (flags 0x%x)", what, h);
  return MO_OK;
}

int edge_store(const float* V, int64_t nv, const int32_t* idx, int64_t n, int32_t pid, int kind, void* stream) {
  Template* t = get_template(pid);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", pid);
  MO_CHECK(V && idx && nv >= 0 && n >= 0, MO_ERR_ARG, "invalid vertex/edge buffers");
  hipStream_t s = as_stream(stream);
  const long long nh = kind == 0 ? 3 * (long long)n : (long long)n;
  int rc;
  if ((rc = t->edge_off.alloc(std::max<long long>(nh, 1) * 3 * sizeof(float)))) return rc;
  Arena ar;
  if ((rc = ar.reserve(256, s))) return rc;
  int* err = ar.take<int>(1);
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  if ((rc = edge_store_launch(V, nv, idx, kind, nh, t->edge_off.as<float>(), err, s))) return rc;
  if ((rc = check_err_flag(err, s, kind == 0 ? "StoreRigidityInformation" : "StoreGraphInformation"))) return rc;
  t->edge_count = nh;
  t->edge_kind = kind;
  return MO_OK;
}

int edge_forward(const float* V, int64_t nv, const int32_t* idx, int64_t n, int32_t pid, float* out, int kind,
                 void* stream) {
  Template* t = get_template(pid);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", pid);
  MO_CHECK(V && idx && out, MO_ERR_ARG, "invalid buffers");
  const long long nh = kind == 0 ? 3 * (long long)n : (long long)n;
  MO_CHECK(t->edge_count >= 0 && t->edge_off.p && nh <= t->edge_count, MO_ERR_STATE,
           "edge loss called with %lld edges but %lld offsets are stored (call the Store function first)", nh,
           t->edge_count);
  hipStream_t s = as_stream(stream);
  // Like the reference (rigid_layer.cc:48-89), every call reads the index
  // buffer as it is now; ends outside [0, nv) raise.
  Arena ar;
  int rc;
  if ((rc = ar.reserve(256, s))) return rc;
  int* err = ar.take<int>(1);
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  if ((rc = edge_fwd_launch(V, nv, idx, kind, nh, t->edge_off.as<float>(), out, err, s))) return rc;
  return check_err_flag(err, s, kind == 0 ? "RigidEdgeLoss_forward" : "GraphEdgeLoss_forward");
}

int edge_backward(const float* V, int64_t nv, const int32_t* idx, int64_t n, int32_t pid, float* out, int kind,
                  void* stream) {
  Template* t = get_template(pid);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", pid);
  MO_CHECK(V && idx && out, MO_ERR_ARG, "invalid buffers");
  const long long nh = kind == 0 ? 3 * (long long)n : (long long)n;
  MO_CHECK(t->edge_off.p && nh <= t->edge_count, MO_ERR_STATE,
           "edge loss backward called with %lld edges but %lld offsets are stored", nh, t->edge_count);
  hipStream_t s = as_stream(stream);
  // The incidence lists are rebuilt from the index buffer on every call
  // (the reference re-reads tensorF each time, rigid_layer.cc:91-133), so a
  // buffer whose contents changed at the same address is never stale.
  const size_t wsb = edge_incidence_workspace_bytes(nh, nv);
  Arena ar;
  int rc;
  if ((rc = ar.reserve(al(wsb) + 256, s))) return rc;
  int* err = ar.take<int>(1);
  void* ws = ar.take<char>(wsb);
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  long long *rp = nullptr, *inc = nullptr;
  if ((rc = edge_incidence_build(idx, kind, nh, nv, ws, &rp, &inc, err, s))) return rc;
  if ((rc = edge_bwd_launch(V, nv, idx, kind, t->edge_off.as<float>(), rp, inc, out, s))) return rc;
  return check_err_flag(err, s, kind == 0 ? "RigidEdgeLoss_backward" : "GraphEdgeLoss_backward");
}

// ---------------------------------------------------------------------------
// Batched LM pipeline.

__global__ void k_stamp(long long* out) { *out = wall_clock64(); }

// x0 = rest position (w: sum of edge weights), u = start - rest (0 when the
// solve starts at the rest shape).
// Per-pair inputs and outputs of the batched set-up / write-back kernels.
struct PairIO {
  const double* x0;       // normalized FP64 source [nv,3]
  const double* xstart;   // start positions, or null (the rest shape)
  const double* sp;       // (scale, pos) for denormalized output, or null
  double* out;
  long long vbase;
  long long nv;
};

// blockIdx.y = pair; x0 = (V, degree), u = xstart - V (or 0).
__global__ void k_make_x0(const PairIO* __restrict__ io, const float* __restrict__ deg, float4* __restrict__ x0,
                          float4* __restrict__ u) {
  const PairIO P = io[blockIdx.y];
  const double* __restrict__ V = P.x0;
  const double* __restrict__ xs = P.xstart;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nv;
       i += (long long)gridDim.x * blockDim.x) {
    const long long g = P.vbase + i;
    x0[g] = make_float4((float)V[i * 3], (float)V[i * 3 + 1], (float)V[i * 3 + 2], deg[g]);
    if (xs)
      u[g] = make_float4((float)(xs[i * 3] - V[i * 3]), (float)(xs[i * 3 + 1] - V[i * 3 + 1]),
                         (float)(xs[i * 3 + 2] - V[i * 3 + 2]), 0.f);
    else
      u[g] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// AdaptiveEdgeLoss weights (edgeloss.h:35-61): every adjacency entry of a
// kind-2 pair gets m * (lambda * 2e-2 / (|V_i - V_j| + 1e-8))^2 (m merged
// directed edges between the two vertices, all with the same rest offset),
// and the vertex's degree becomes the sum of its weights.  V is the FP64
// rest shape; adj / deg are this pair's slices.
__global__ void k_adaptive_weights(const double* __restrict__ V, int nv, const int2* __restrict__ adj,
                                   const unsigned* __restrict__ nbr, double lambda, float* __restrict__ wt,
                                   float* __restrict__ deg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int2 a = adj[i];
  double dsum = 0.0;
  for (int k = 0; k < a.y; ++k) {
    const unsigned en = nbr[a.x + k];
    const int j = (int)(en & 0xFFFFFFu);
    const double m = (double)(en >> 24);
    const double d0 = V[i * 3] - V[j * 3], d1 = V[i * 3 + 1] - V[j * 3 + 1], d2 = V[i * 3 + 2] - V[j * 3 + 2];
    const double l = lambda * (2e-2 / (sqrt(d0 * d0 + d1 * d1 + d2 * d2) + 1e-8));
    const double w = m * (l * l);
    wt[a.x + k] = (float)w;
    dsum += w;
  }
  deg[i] = (float)dsum;
}

// out = x0_64 + u (normalized frame), or (x0_64 + u) * scale + pos when sp
// is given (Mesh::WriteOBJ, mesh.cc:46-64).
// blockIdx.y = pair: out = denormalized (x0 + u).
__global__ void k_finish(const PairIO* __restrict__ io, const float4* __restrict__ u) {
  const PairIO P = io[blockIdx.y];
  const double* __restrict__ X = P.x0;
  const double* __restrict__ sp = P.sp;
  double* __restrict__ out = P.out;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nv;
       i += (long long)gridDim.x * blockDim.x) {
    const float4 d = u[P.vbase + i];
    double p[3] = {X[i * 3] + (double)d.x, X[i * 3 + 1] + (double)d.y, X[i * 3 + 2] + (double)d.z};
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = sp ? p[k] * sp[0] + sp[1 + k] : p[k];
  }
}

struct CorePair {
  const float* grid;
  int N;
  const double* x0;    // normalized FP64 source [nv,3]
  long long nv;
  const int32_t* idx;
  long long nidx;
  int kind;
  double lambda;
  const double* sp;    // device (scale, pos) for denormalized output, or null
  double* out;
  const double* xstart = nullptr;  // start positions (device FP64), null = the rest shape x0
};

LMOptionsDev to_dev_options(const mo_lm_options& o) {
  LMOptionsDev d;
  d.max_num_iterations = o.max_num_iterations;
  d.max_consecutive_invalid_steps = o.max_consecutive_invalid_steps;
  d.function_tolerance = o.function_tolerance;
  d.gradient_tolerance = o.gradient_tolerance;
  d.parameter_tolerance = o.parameter_tolerance;
  d.min_relative_decrease = o.min_relative_decrease;
  d.initial_trust_region_radius = o.initial_trust_region_radius;
  d.max_trust_region_radius = o.max_trust_region_radius;
  d.min_trust_region_radius = o.min_trust_region_radius;
  d.min_lm_diagonal = o.min_lm_diagonal;
  d.max_lm_diagonal = o.max_lm_diagonal;
  d.pcg_max_iterations = o.pcg_max_iterations;
  d.pcg_relative_tolerance = o.pcg_relative_tolerance;
  d.max_successful_steps = o.max_successful_steps;
  return d;
}

// HIP events around the phases of a batch call (only when opt.timing is set).
struct PhaseTimer {
  bool on = false;
  bool setup_first = false;   // marks 1 and 2 close the set-up and then the grid span
  hipStream_t s = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  long long* stamp = nullptr;   // device: wall clock at the start of the call
  int start(bool enable, hipStream_t st) {
    on = enable;
    s = st;
    if (!on) return MO_OK;
    for (auto& e : ev) MO_HIP(hipEventCreate(&e));
    MO_HIP(hipMallocAsync((void**)&stamp, sizeof(long long), s));
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, stamp);
    MO_HIP(hipEventRecord(ev[0], s));
    return MO_OK;
  }
  int mark(int i) {
    if (on) MO_HIP(hipEventRecord(ev[i], s));
    return MO_OK;
  }
  int finish() {
    if (!on) return MO_OK;
    MO_HIP(hipEventRecord(ev[4], s));
    MO_HIP(hipEventSynchronize(ev[4]));
    float ms[4];
    MO_HIP(hipEventElapsedTime(&ms[0], ev[0], ev[1]));
    MO_HIP(hipEventElapsedTime(&ms[1], ev[1], ev[2]));
    MO_HIP(hipEventElapsedTime(&ms[2], ev[2], ev[3]));
    MO_HIP(hipEventElapsedTime(&ms[3], ev[0], ev[4]));
    if (setup_first) std::swap(ms[0], ms[1]);
    long long st = 0;
    MO_HIP(hipMemcpy(&st, stamp, sizeof(long long), hipMemcpyDeviceToHost));
    int dev = 0, rate_khz = 0;
    MO_HIP(hipGetDevice(&dev));
    MO_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    for (int k = 0; k < 4; ++k) g_timings[k] = ms[k];
    g_timings[4] = (double)st;
    g_timings[5] = (double)rate_khz * 1000.0;
    return MO_OK;
  }
  ~PhaseTimer() {
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    if (stamp) (void)hipFreeAsync(stamp, s);
  }
};

int device_lds_limit() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
      cached = v;
    else
      cached = 65536;
  }
  return cached;
}

void fill_stats(const LMResultDev* res, const int* path, const int* wgs, int np, mo_lm_stats* stats) {
  for (int p = 0; p < np; ++p) {
    const LMResultDev& r = res[p];
    mo_lm_stats& o = stats[p];
    o.initial_cost = r.initial_cost;
    o.final_cost = r.final_cost;
    o.vertex_cost = r.vertex_cost;
    o.rigidity_cost = r.rigidity_cost;
    o.iterations = r.iterations;
    o.successful_steps = r.successful_steps;
    o.unsuccessful_steps = r.unsuccessful_steps;
    o.termination = r.termination;
    o.pcg_iterations = r.pcg_iterations;
    o.clock_start = r.clock_start;
    o.clock_end = r.clock_end;
    o.solver_path = path[p];
    o.workgroups = wgs[p];
  }
}

// ext / ext_cap: a caller-provided device region for all of the solve's
// scratch (the no-allocation entry point); null: stream-ordered allocations.
constexpr int kMaxLdsLaunches = 32;   // register/LDS launches (shapes) per batch

int lm_core(const std::vector<CorePair>& pairs, const mo_lm_options& opt, mo_lm_stats* stats, hipStream_t s,
            PhaseTimer& timer, char* ext = nullptr, size_t ext_cap = 0,
            const std::function<int()>& between = {}) {
  const int np = (int)pairs.size();
  if (np == 0) return MO_OK;
  MO_CHECK(opt.block_threads >= 64 && opt.block_threads <= 1024 && opt.block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 1024]");
  long long nvtot = 0, nedges_tot = 0, max_edges = 0, max_nv = 0;
  std::vector<long long> vbase(np);
  for (int p = 0; p < np; ++p) {
    const CorePair& c = pairs[p];
    MO_CHECK(c.nv >= 0 && c.nv < (1 << 24), MO_ERR_ARG, "pair %d: vertex count %lld outside [0, 2^24)", p, c.nv);
    MO_CHECK(c.kind >= 0 && c.kind <= 2, MO_ERR_ARG,
             "pair %d: kind must be 0 (faces), 1 (edges) or 2 (edges, adaptive weights)", p);
    vbase[p] = nvtot;
    nvtot += c.nv;
    long long ne = c.kind == 0 ? 3 * c.nidx : c.nidx;
    nedges_tot += ne;
    max_edges = std::max(max_edges, ne);
    max_nv = std::max(max_nv, c.nv);
  }
  const long long nvalloc = std::max<long long>(nvtot, 1);
  size_t bytes = al(nvalloc * sizeof(float4)) * 2 + al(nvalloc * sizeof(int2)) + al(nvalloc * sizeof(float)) +
                 al((size_t)kLMWorkSlots * nvalloc * sizeof(float4)) +
                 al(std::max<long long>(2 * nedges_tot, 1) * sizeof(unsigned)) +
                 al(std::max<long long>(2 * nedges_tot, 1) * sizeof(float)) + al(topo_workspace_bytes(nvalloc)) +
                 al(np * sizeof(LMPairDesc)) + al(np * sizeof(TopoPairDesc)) + al(np * sizeof(LMResultDev)) +
                 al(2 * np * sizeof(int)) + al(np * sizeof(int)) + al(4 * np * sizeof(int)) +
                 al(np * sizeof(PairIO)) + al(kMaxLdsLaunches * sizeof(int)) + 512;
  MO_CHECK(np <= 65535, MO_ERR_ARG, "at most 65535 pairs per batch (got %d)", np);
  Arena ar;
  int rc;
  if ((rc = ext ? ar.attach(ext, ext_cap, bytes, s) : ar.reserve(bytes, s))) return rc;
  float4* x0 = ar.take<float4>(nvalloc);
  float4* u = ar.take<float4>(nvalloc);
  int2* adj = ar.take<int2>(nvalloc);
  float* deg = ar.take<float>(nvalloc);
  float4* work = ar.take<float4>((size_t)kLMWorkSlots * nvalloc);
  unsigned* raw = ar.take<unsigned>(std::max<long long>(2 * nedges_tot, 1));
  float* wt = ar.take<float>(std::max<long long>(2 * nedges_tot, 1));
  void* tws = ar.take<char>(topo_workspace_bytes(nvalloc));
  LMPairDesc* d_pairs = ar.take<LMPairDesc>(np);
  TopoPairDesc* d_topo = ar.take<TopoPairDesc>(np);
  LMResultDev* d_res = ar.take<LMResultDev>(np);
  int* d_pstat = ar.take<int>(2 * np);
  int* d_plist = ar.take<int>(np);
  int* d_chunk = ar.take<int>(4 * np);
  PairIO* d_io = ar.take<PairIO>(np);
  int* d_next = ar.take<int>(kMaxLdsLaunches);   // work counters of the register/LDS launches
  int* err = ar.take<int>(1);

  std::vector<LMPairDesc> h_pairs(np);
  std::vector<TopoPairDesc> h_topo(np);
  std::vector<PairIO> h_io(np);
  long long max_nv_io = 1;
  for (int p = 0; p < np; ++p) {
    const CorePair& c = pairs[p];
    h_io[p] = PairIO{c.x0, c.xstart, c.sp, c.out, vbase[p], c.nv};
    max_nv_io = std::max(max_nv_io, c.nv);
    const bool weighted = c.kind == 2;
    h_pairs[p] = LMPairDesc{c.grid, c.N, (int)c.nv, vbase[p], 0, weighted ? 1.0f : (float)c.lambda, weighted ? 1 : 0};
    h_topo[p] = TopoPairDesc{c.idx, c.kind == 0 ? 3 * c.nidx : c.nidx, vbase[p], (int)c.nv, c.kind == 0 ? 0 : 1};
  }
  // With work to enqueue in between (a batch's grid builds), the topology
  // runs on a second stream, forked from s here and joined back into s
  // after its sizes are read, so that the two overlap (not inside a batch
  // graph's record / capture, which keep one stream, and not with phase
  // timing on, whose events time the phases one after the other).
  AuxStream aux;
  struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
      if (e) (void)hipEventDestroy(e);
    }
  } fork_ev, join_ev;
  hipStream_t ts = s;
  if (between && !g_chain && opt.timing == 0) {
    if ((rc = aux.acquire())) return rc;
    MO_HIP(hipEventCreateWithFlags(&fork_ev.e, hipEventDisableTiming));
    MO_HIP(hipEventCreateWithFlags(&join_ev.e, hipEventDisableTiming));
    MO_HIP(hipEventRecord(fork_ev.e, s));
    MO_HIP(hipStreamWaitEvent(aux.s, fork_ev.e, 0));
    ts = aux.s;
  }
  if ((rc = chain_h2d(d_pairs, h_pairs.data(), np * sizeof(LMPairDesc), ts))) return rc;
  if ((rc = chain_h2d(d_topo, h_topo.data(), np * sizeof(TopoPairDesc), ts))) return rc;
  if ((rc = chain_h2d(d_io, h_io.data(), np * sizeof(PairIO), ts))) return rc;
  if ((rc = zero_fill(err, sizeof(int), ts))) return rc;
  if ((rc = topo_build_launch(d_topo, np, max_edges, max_nv, nvtot, raw, adj, deg, d_pstat, tws, err, ts))) return rc;
  for (int p = 0; p < np; ++p) {
    const CorePair& c = pairs[p];
    if (c.nv == 0 || c.kind != 2) continue;
    hipLaunchKernelGGL(k_adaptive_weights, dim3((unsigned)((c.nv + 255) / 256)), dim3(256), 0, ts, c.x0, (int)c.nv,
                       adj + vbase[p], raw, c.lambda, wt, deg + vbase[p]);
  }
  const dim3 io_grid((unsigned)std::min<long long>((max_nv_io + 255) / 256, 64), (unsigned)np);
  hipLaunchKernelGGL(k_make_x0, io_grid, dim3(256), 0, ts, d_io, deg, x0, u);
  MO_HIP(hipGetLastError());
  // per pair and register/LDS team shape: the largest chunk's adjacency entries
  if ((rc = lm_lds_chunk_entries(d_pairs, adj, np, d_chunk, ts))) return rc;
  std::vector<int> pstat(2 * np), hchunk(4 * np);
  int herr = 0;
  {
    const ChainRead rd[3] = {{hchunk.data(), d_chunk, 4 * np * sizeof(int)},
                             {pstat.data(), d_pstat, 2 * np * sizeof(int)},
                             {&herr, err, sizeof(int)}};
    if ((rc = chain_read(rd, 3, ts, between))) return rc;
  }
  if (ts != s) {
    MO_HIP(hipEventRecord(join_ev.e, ts));
    MO_HIP(hipStreamWaitEvent(s, join_ev.e, 0));
  }
  MO_CHECK(herr == 0, MO_ERR_RANGE, "source topology: index out of range or too many vertices (flags 0x%x)", herr);

  // Schedule: largest pairs first within every launch (workgroups are
  // dispatched in order, so this is longest-processing-time-first).  Pairs
  // that fit the register/LDS kernel go to its size buckets, the rest to the
  // global-memory kernel.
  std::vector<int> order(np);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pairs[a].nv > pairs[b].nv; });
  const int nb = lm_lds_buckets();
  const size_t lds_cap = (size_t)std::max(0, device_lds_limit() - 4096);
  std::vector<std::vector<int>> bucket(nb + 1);
  std::vector<size_t> bucket_lds(nb + 1, 0);
  for (int p : order) {
    const CorePair& c = pairs[p];
    int b = -1;
    // the register/LDS kernel packs a vertex's LDS adjacency offset and count
    // into 16 bits each and a neighbour index (or the zero slot, index nv)
    // into 15: a pair qualifies only when all of them fit
    const int padded = hchunk[4 * p + 3];   // entries with every count rounded up to even
    if (opt.kernel_variant == 0 && c.kind != 2 && c.nv > 0 && c.nv <= 32767 && pstat[2 * p + 1] <= 2 &&
        padded < 65536) {
      b = lm_lds_bucket_for((int)c.nv);
      if (b >= 0 && lm_lds_bytes((int)c.nv, padded) > lds_cap) b = -1;
    }
    if (b < 0) b = nb;
    bucket[b].push_back(p);
    if (b < nb) bucket_lds[b] = std::max(bucket_lds[b], lm_lds_bytes((int)c.nv, padded));
  }
  std::vector<int> plist;
  std::vector<int> start(nb + 2, 0);
  for (int b = 0; b <= nb; ++b) {
    start[b] = (int)plist.size();
    plist.insert(plist.end(), bucket[b].begin(), bucket[b].end());
  }
  start[nb + 1] = (int)plist.size();
  if ((rc = chain_h2d(d_plist, plist.data(), np * sizeof(int), s))) return rc;
  LMArrays A{x0, u, adj, raw, work, wt};
  const LMOptionsDev dopt = to_dev_options(opt);
  MO_CHECK(nb <= kMaxLdsLaunches, MO_ERR_STATE, "%d register/LDS shapes exceed the %d work counters", nb,
           kMaxLdsLaunches);
  if ((rc = zero_fill(d_next, nb * sizeof(int), s))) return rc;
  if (AsyncCtx* a = g_async) {
    if (!a->lm_ready) MO_HIP(hipEventCreateWithFlags(&a->lm_ready, hipEventDisableTiming));
    MO_HIP(hipEventRecord(a->lm_ready, s));
  }
  if ((rc = timer.mark(2))) return rc;
  for (int b = 0; b < nb; ++b)
    if (!bucket[b].empty())
      if ((rc = lm_lds_launch(b, d_pairs, d_plist + start[b], (int)bucket[b].size(), d_next + b, A, dopt, d_res,
                              nvalloc, bucket_lds[b], s)))
        return rc;
  // Pairs the register/LDS kernel does not take.  A pair of nv vertices
  // gets a team of ceil(nv / team_vertices) co-resident workgroups (at most
  // lm_team_max(); team_vertices 0 means 4 block_threads, or block_threads
  // under kernel_variant 2, which sends every pair here), a function of the
  // pair and the options alone so its result does not depend on the rest of
  // the batch.  Teams of one run on k_lm_solve
  // (kernel_variant 2 keeps them on the team kernel); the others are packed,
  // largest first, into cooperative launches of at most lm_team_capacity()
  // workgroups.  One team launch at a time in the process (team_launch_mutex, held
  // until the final stream synchronization): two cooperative grids from
  // different streams must not wait on each other's residency.
  std::vector<int> solo, teams, team_c;
  std::vector<int4> wg;
  std::vector<int> path(np, MO_PATH_LDS);
  std::vector<int> nwgs(np, 1);   // workgroups that solve each pair
  for (int p : bucket[nb]) path[p] = MO_PATH_ONE_WORKGROUP;
  std::unique_lock<std::mutex> team_lock(team_launch_mutex(), std::defer_lock);
  // leaving with the lock held (an error after a cooperative launch): wait
  // for the grid first, so the next team launch cannot overlap it
  struct TeamDrain {
    std::unique_lock<std::mutex>& lk;
    hipStream_t s;
    ~TeamDrain() {
      if (lk.owns_lock()) (void)hipStreamSynchronize(s);
    }
  } team_drain{team_lock, s};
  int* d_abort = nullptr;
  Arena tar;
  {
    const int bt = opt.block_threads;
    MO_CHECK(opt.team_vertices >= 0, MO_ERR_ARG, "team_vertices must be >= 0");
    // Pairs above one register/LDS workgroup go to the register/LDS team
    // kernel when their multiplicities allow it (kind 0/1, <= 2), with the
    // first team shape in preference order (lm_lds_team_order) whose chunk
    // fits the LDS, a team of at most kMaxTeam workgroups, and the whole
    // team co-resident.  The rest take the global-memory kernels below.
    std::vector<int> lds_cand;
    std::vector<int> rest;
    // a pair that no team shape can hold alone (its team would exceed
    // kMaxTeam workgroups, a chunk's entries the 16-bit LDS offsets, or its
    // LDS the budget) takes the global-memory kernels
    auto lds_team_fits = [&](int p, int b) {
      int T = 0, V = 0;
      if (lm_lds_team_shape(b, &T, &V)) return false;
      const long long chcap = (long long)T * V, nvp = pairs[p].nv;
      const long long c = (nvp + chcap - 1) / chcap, ch = (nvp + c - 1) / c;
      const long long ent = hchunk[4 * p + b];
      return c <= lm_lds_team_max() && ent < 65536 && lm_lds_team_bytes((int)ch, ent) <= lds_cap;
    };
    for (int p : bucket[nb]) {
      const CorePair& c = pairs[p];
      bool ok = opt.kernel_variant == 0 && c.kind != 2 && c.nv > 0 && pstat[2 * p + 1] <= 2;
      if (ok) {
        ok = false;
        for (int b : lm_lds_team_order()) ok = ok || lds_team_fits(p, b);
      }
      (ok ? lds_cand : rest).push_back(p);
    }
    // each pair takes the first shape in preference order that holds it, so its
    // chunking -- and its result -- does not depend on the rest of the
    // batch; the teams are grouped by shape, one cooperative grid per
    // group, its LDS size the group's largest
    const int nls = lm_lds_team_shapes();
    std::vector<int> pshape(lds_cand.size(), -1);
    std::vector<size_t> lds_of(nls, 0);
    for (size_t k = 0; k < lds_cand.size(); ++k) {
      const int p = lds_cand[k];
      for (int b : lm_lds_team_order())
        if (pshape[k] < 0 && lds_team_fits(p, b)) {
          int T = 0, V = 0;
          if ((rc = lm_lds_team_shape(b, &T, &V))) return rc;
          const long long chcap = (long long)T * V, nvp = pairs[p].nv;
          const long long c = (nvp + chcap - 1) / chcap, ch = (nvp + c - 1) / c;
          pshape[k] = b;
          lds_of[b] = std::max(lds_of[b], lm_lds_team_bytes((int)ch, hchunk[4 * p + b]));
        }
    }
    std::vector<int> cap_of(nls, 0);
    for (int b = 0; b < nls; ++b)
      if (lds_of[b] > 0) cap_of[b] = lm_lds_team_capacity(b, lds_of[b]);
    std::vector<int> lteams, lteam_c, lteam_ch, lteam_b;
    for (int b = 0; b < nls; ++b) {
      int T = 0, V = 0;
      if ((rc = lm_lds_team_shape(b, &T, &V))) return rc;
      const long long chcap = (long long)T * V;
      for (size_t k = 0; k < lds_cand.size(); ++k) {
        if (pshape[k] != b) continue;
        const int p = lds_cand[k];
        const long long nvp = pairs[p].nv;
        const long long c = (nvp + chcap - 1) / chcap;
        if (c > cap_of[b]) {   // the team cannot be co-resident
          rest.push_back(p);
          continue;
        }
        lteams.push_back(p);
        lteam_c.push_back((int)c);
        lteam_ch.push_back((int)((nvp + c - 1) / c));
        lteam_b.push_back(b);
        path[p] = MO_PATH_LDS_TEAM;
        nwgs[p] = (int)c;
      }
    }
    const long long vpw = opt.team_vertices > 0 ? opt.team_vertices : (opt.kernel_variant == 2 ? bt : 4LL * bt);
    const int cap = opt.kernel_variant == 1 ? 0 : lm_team_capacity(bt);
    for (int p : rest) {
      const long long c = std::min<long long>(std::max<long long>((pairs[p].nv + vpw - 1) / vpw, 1), lm_team_max());
      if (cap >= lm_team_max() && (c > 1 || opt.kernel_variant == 2)) {
        teams.push_back(p);
        team_c.push_back((int)c);
        path[p] = MO_PATH_TEAM;
        nwgs[p] = (int)c;
      } else {
        solo.push_back(p);
      }
    }
    if (!solo.empty()) {
      if ((rc = chain_h2d(d_plist + start[nb], solo.data(), solo.size() * sizeof(int), s))) return rc;
      if ((rc = lm_solve_launch(d_pairs, d_plist + start[nb], (int)solo.size(), A, dopt, d_res, nvalloc, bt, s)))
        return rc;
    }
    // team arrays: the register/LDS teams first, then the global-memory teams
    const int nlt = (int)lteams.size(), nt = nlt + (int)teams.size();
    if (nt > 0 && g_chain) {
      // cooperative grids are not captured into batch graphs
      MO_CHECK(g_chain->mode == ChainCtx::kRecord, MO_ERR_STATE, "batch graph capture: unexpected team launch");
      g_chain->team_launch = true;
    }
    if (nt > 0) {
      std::vector<int> all_t(lteams);
      all_t.insert(all_t.end(), teams.begin(), teams.end());
      const int per = lm_team_part_doubles();
      long long nwg_all = 0;
      for (int c : lteam_c) nwg_all += c;
      for (int c : team_c) nwg_all += c;
      const size_t team_bytes = al(nt * sizeof(int)) + al(nwg_all * sizeof(int4)) +
                                al((size_t)nt * 32 * sizeof(unsigned) + sizeof(int)) +
                                al((size_t)nt * per * sizeof(double)) + 512;
      if ((rc = ext ? tar.attach(ext + ar.end(), ext_cap - std::min(ext_cap, ar.end()), team_bytes, s)
                    : tar.reserve(team_bytes, s)))
        return rc;
      int* d_tp = tar.take<int>(nt);
      int4* d_wg = tar.take<int4>(nwg_all);
      unsigned* d_cnt = tar.take<unsigned>((size_t)nt * 32 + 1);
      double* d_part = tar.take<double>((size_t)nt * per);
      d_abort = reinterpret_cast<int*>(d_cnt + (size_t)nt * 32);
      if ((rc = zero_fill(d_cnt, ((size_t)nt * 32 + 1) * sizeof(unsigned), s))) return rc;
      // chunks: consecutive teams (largest first) whose workgroups fit one
      // grid; a chunk never mixes the two kernels
      std::vector<int> chunk_team{0}, chunk_wg{0};
      int used = 0;
      for (int t = 0; t < nt; ++t) {
        const bool lt = t < nlt;
        const int c = lt ? lteam_c[t] : team_c[t - nlt];
        if (t > 0 && (used + c > (lt ? cap_of[lteam_b[t]] : cap) || t == nlt || (lt && lteam_b[t] != lteam_b[t - 1]))) {
          chunk_team.push_back(t);
          chunk_wg.push_back((int)wg.size());
          used = 0;
        }
        const int rel = t - chunk_team.back();
        for (int r = 0; r < c; ++r) wg.push_back(make_int4(rel, r, c, lt ? lteam_ch[t] : 0));
        used += c;
      }
      chunk_team.push_back(nt);
      chunk_wg.push_back((int)wg.size());
      MO_HIP(hipMemcpyAsync(d_tp, all_t.data(), nt * sizeof(int), hipMemcpyHostToDevice, s));
      MO_HIP(hipMemcpyAsync(d_wg, wg.data(), wg.size() * sizeof(int4), hipMemcpyHostToDevice, s));
      team_lock.lock();
      for (size_t k = 0; k + 1 < chunk_team.size(); ++k) {
        const int t0 = chunk_team[k], t1 = chunk_team[k + 1];
        const bool lt = t0 < nlt;
        if (lt)
          rc = lm_lds_team_launch(lteam_b[t0], d_pairs, d_tp + t0, d_wg + chunk_wg[k], chunk_wg[k + 1] - chunk_wg[k],
                                  A, dopt, d_res, nvalloc, lds_of[lteam_b[t0]], d_cnt + (size_t)t0 * 32,
                                  d_part + (size_t)t0 * per, d_abort, s);
        else
          rc = lm_team_launch(d_pairs, d_tp + t0, d_wg + chunk_wg[k], chunk_wg[k + 1] - chunk_wg[k], A, dopt, d_res,
                              nvalloc, bt, d_cnt + (size_t)t0 * 32, d_part + (size_t)t0 * per, d_abort, s);
        if (rc == kCoopRefused) {
          // the runtime refused the cooperative grid: this chunk's and the
          // remaining teams' pairs run one workgroup each (d_tp lists them)
          if ((rc = lm_solve_launch(d_pairs, d_tp + t0, nt - t0, A, dopt, d_res, nvalloc, bt, s))) return rc;
          for (int t = t0; t < nt; ++t) {
            path[all_t[t]] = MO_PATH_COOP_REFUSED;
            nwgs[all_t[t]] = 1;
          }
          break;
        }
        if (rc) return rc;
        (void)t1;
      }
    }
  }
  if ((rc = timer.mark(3))) return rc;
  if (d_abort) {
    // a team whose barrier wait ran out leaves invalid solutions: check
    // before anything is written to the caller's buffers
    int h_abort = 0;
    MO_HIP(hipMemcpyAsync(&h_abort, d_abort, sizeof(int), hipMemcpyDeviceToHost, s));
    MO_HIP(hipStreamSynchronize(s));
    MO_CHECK(h_abort == 0, MO_ERR_HIP,
             "LM team solve: a barrier wait exceeded its bound (results invalid; outputs not written)");
  }
  hipLaunchKernelGGL(k_finish, io_grid, dim3(256), 0, s, d_io, u);
  MO_HIP(hipGetLastError());
  if (ChainCtx* c = g_chain) {
    if (c->mode == ChainCtx::kCapture) {
      // the graph ends with the results in its pinned buffer; stats are
      // filled after each replay
      size_t off = 0;
      if ((rc = pinned_take(c, np * sizeof(LMResultDev), &off))) return rc;
      MO_HIP(hipMemcpyAsync(c->pinned + off, d_res, np * sizeof(LMResultDev), hipMemcpyDeviceToHost, s));
      c->res_off = off;
      c->res_np = np;
      c->res_path = path;
      c->res_wgs = nwgs;
      return MO_OK;
    }
    c->pinned_bytes += al256c(np * sizeof(LMResultDev));
  }
  if (AsyncCtx* a = g_async) {
    if (!d_abort) {
      MO_HIP(hipMemcpyAsync(a->pinned, d_res, np * sizeof(LMResultDev), hipMemcpyDeviceToHost, s));
      a->deferred = true;
      a->np = np;
      a->path = path;
      a->wgs = nwgs;
      return MO_OK;
    }
  }
  std::vector<LMResultDev> h_res(np);
  MO_HIP(hipMemcpyAsync(h_res.data(), d_res, np * sizeof(LMResultDev), hipMemcpyDeviceToHost, s));
  MO_HIP(hipStreamSynchronize(s));
  if (stats) fill_stats(h_res.data(), path.data(), nwgs.data(), np, stats);
  return MO_OK;
}

}  // namespace
}  // namespace mo

using namespace mo;

// ===========================================================================
// Exported entry points.

MO_EXPORT int mo_version(void) { return 100; }

MO_EXPORT const char* mo_last_error(void) { return g_last_error.c_str(); }

MO_EXPORT int mo_template_create_f32(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t symmetry,
                                     int32_t grid_resolution, void* stream, int32_t* param_id) {
  (void)symmetry;   // no-op in the reference (deform_params.cc:28-29)
  return template_create<float>(V, nv, F, nf, grid_resolution, stream, param_id);
}

MO_EXPORT int mo_template_create_f64(const double* V, int64_t nv, const int32_t* F, int64_t nf, int32_t symmetry,
                                     int32_t grid_resolution, void* stream, int32_t* param_id) {
  (void)symmetry;
  return template_create<double>(V, nv, F, nf, grid_resolution, stream, param_id);
}

MO_EXPORT int mo_template_rebuild_f64(int32_t param_id, const double* V, int64_t nv, const int32_t* F, int64_t nf,
                                      void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  int dev = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_CHECK(dev == t->device, MO_ERR_ARG, "template %d lives on device %d, not the current device %d", param_id,
           t->device, dev);
  t->edge_off.reset();
  t->edge_count = 0;
  t->edge_kind = -1;
  const int rc = template_fill<double>(t, V, nv, F, nf, as_stream(stream));
  if (rc) t->scale = 0.0;   // a failed rebuild leaves no usable template behind
  return rc;
}

MO_EXPORT int mo_template_destroy(int32_t param_id) {
  std::lock_guard<std::mutex> lk(g_mu);
  MO_CHECK(param_id >= 0 && param_id < (int)g_templates.size() && g_templates[param_id], MO_ERR_ARG,
           "unknown param_id %d", param_id);
  g_templates[param_id].reset();
  return MO_OK;
}

MO_EXPORT int mo_template_info(int32_t param_id, int32_t* grid_resolution, double* scale, double* trans3) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  if (grid_resolution) *grid_resolution = t->N;
  if (scale) *scale = t->scale;
  if (trans3) for (int k = 0; k < 3; ++k) trans3[k] = t->trans[k];
  return MO_OK;
}

MO_EXPORT int mo_template_grid(int32_t param_id, double* dist64, float* dist32, int32_t* nearest, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  hipStream_t s = as_stream(stream);
  if (dist64) MO_HIP(hipMemcpyAsync(dist64, t->grid64.p, t->cells * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (dist32) MO_HIP(hipMemcpyAsync(dist32, t->grid32.p, t->cells * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (nearest) MO_HIP(hipMemcpyAsync(nearest, t->nearest.p, t->cells * sizeof(int), hipMemcpyDeviceToDevice, s));
  return MO_OK;
}

MO_EXPORT int mo_grid_build(const double* Vn, int64_t nv, const int32_t* F, int64_t nf, int32_t grid_resolution,
                            double* dist64, float* dist32, int32_t* nearest, void* stream) {
  MO_CHECK(Vn && F && nv > 0 && nf > 0, MO_ERR_ARG, "grid build needs vertices and faces");
  hipStream_t s = as_stream(stream);
  GridSpec gs{Vn, nv, F, nf, grid_resolution, dist64, dist32, nearest};
  const size_t gws = grid_batch_workspace_bytes(&gs, 1);
  Arena ar;
  int rc;
  if ((rc = ar.reserve(al(gws) + 256, s))) return rc;
  void* ws = ar.take<char>(gws);
  int* err = ar.take<int>(1);
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  if ((rc = grid_build_batch_launch(&gs, 1, ws, s, err))) return rc;
  return check_err_flag(err, s, "mo_grid_build");
}

MO_EXPORT int mo_normalize_by_template(float* V, int64_t nv, int32_t param_id, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  MO_CHECK(V || nv == 0, MO_ERR_ARG, "null vertex buffer");
  return normalize_by_template_launch(V, nv, t->scale, t->trans, false, as_stream(stream));
}

MO_EXPORT int mo_denormalize_by_template(float* V, int64_t nv, int32_t param_id, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  MO_CHECK(V || nv == 0, MO_ERR_ARG, "null vertex buffer");
  return normalize_by_template_launch(V, nv, t->scale, t->trans, true, as_stream(stream));
}

MO_EXPORT int mo_distance_loss_forward(const float* V, int64_t nv, int32_t param_id, float* out, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  MO_CHECK((V && out) || nv == 0, MO_ERR_ARG, "null buffer");
  return distance_loss_launch(t->grid32.as<float>(), t->N, V, nv, out, false, as_stream(stream));
}

MO_EXPORT int mo_distance_loss_backward(const float* V, int64_t nv, int32_t param_id, float* out, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  MO_CHECK((V && out) || nv == 0, MO_ERR_ARG, "null buffer");
  return distance_loss_launch(t->grid32.as<float>(), t->N, V, nv, out, true, as_stream(stream));
}

MO_EXPORT int mo_rigidity_store(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id,
                                void* stream) {
  return edge_store(V, nv, F, nf, param_id, 0, stream);
}

MO_EXPORT int mo_rigid_loss_forward(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id,
                                    float* out, void* stream) {
  return edge_forward(V, nv, F, nf, param_id, out, 0, stream);
}

MO_EXPORT int mo_rigid_loss_backward(const float* V, int64_t nv, const int32_t* F, int64_t nf, int32_t param_id,
                                     float* out, void* stream) {
  return edge_backward(V, nv, F, nf, param_id, out, 0, stream);
}

MO_EXPORT int mo_graph_store(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id,
                             void* stream) {
  return edge_store(V, nv, E, ne, param_id, 1, stream);
}

MO_EXPORT int mo_graph_loss_forward(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id,
                                    float* out, void* stream) {
  return edge_forward(V, nv, E, ne, param_id, out, 1, stream);
}

MO_EXPORT int mo_graph_loss_backward(const float* V, int64_t nv, const int32_t* E, int64_t ne, int32_t param_id,
                                     float* out, void* stream) {
  return edge_backward(V, nv, E, ne, param_id, out, 1, stream);
}

MO_EXPORT void mo_lm_default_options(mo_lm_options* o) {
  if (!o) return;
  o->max_num_iterations = 100;
  o->max_consecutive_invalid_steps = 5;
  o->function_tolerance = 1e-6;
  o->gradient_tolerance = 1e-10;
  o->parameter_tolerance = 1e-8;
  o->min_relative_decrease = 1e-3;
  o->initial_trust_region_radius = 1e4;
  o->max_trust_region_radius = 1e16;
  o->min_trust_region_radius = 1e-32;
  o->min_lm_diagonal = 1e-6;
  o->max_lm_diagonal = 1e32;
  o->pcg_max_iterations = 1000;
  o->pcg_relative_tolerance = 1e-5f;
  o->block_threads = 1024;
  o->timing = 0;
  o->kernel_variant = 0;
  o->max_successful_steps = 0;
  o->team_vertices = 0;
}

MO_EXPORT int32_t mo_lm_shape_count(void) { return lm_lds_buckets(); }

MO_EXPORT int mo_lm_shape_info(int32_t index, int32_t* threads, int32_t* vertices_per_thread) {
  int t = 0, v = 0;
  const int rc = lm_lds_shape(index, &t, &v);
  if (rc) return rc;
  if (threads) *threads = t;
  if (vertices_per_thread) *vertices_per_thread = v;
  return MO_OK;
}

MO_EXPORT int mo_lm_set_shape_table(const int32_t* max_vertices, const int32_t* shape, int32_t n) {
  return lm_lds_set_table(max_vertices, shape, n);
}

MO_EXPORT int mo_last_grid_stats(int64_t* out4) {
  MO_CHECK(out4, MO_ERR_ARG, "null output");
  for (int k = 0; k < kGridStatCount; ++k) out4[k] = g_grid_stats[k];
  return MO_OK;
}

MO_EXPORT int mo_last_timings(double* out6) {
  MO_CHECK(out6, MO_ERR_ARG, "null output");
  for (int k = 0; k < 6; ++k) out6[k] = g_timings[k];
  return MO_OK;
}

namespace {

int deform_batch_impl(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                      const int32_t* const* idx, const int64_t* nidx, const int32_t* kind, const double* lambda,
                      const mo_lm_options* opt, mo_lm_stats* stats, void* stream, char* ext, size_t ext_cap) {
  MO_CHECK(npairs >= 0, MO_ERR_ARG, "npairs < 0");
  if (npairs == 0) return MO_OK;
  MO_CHECK(param_ids && V && nv && idx && nidx && kind && lambda, MO_ERR_ARG, "null per-pair array");
  mo_lm_options o;
  if (opt) o = *opt; else mo_lm_default_options(&o);
  MO_CHECK(!ext || o.timing == 0, MO_ERR_ARG,
           "mo_deform_batch_noalloc: options.timing needs allocations; set it to 0");
  std::vector<CorePair> pairs(npairs);
  for (int p = 0; p < npairs; ++p) {
    Template* t = get_template(param_ids[p]);
    MO_CHECK(t, MO_ERR_ARG, "pair %d: unknown param_id %d", p, param_ids[p]);
    MO_CHECK(V[p] || nv[p] == 0, MO_ERR_ARG, "pair %d: null vertex buffer", p);
    MO_CHECK(idx[p] || nidx[p] == 0, MO_ERR_ARG, "pair %d: null index buffer", p);
    pairs[p] = CorePair{t->grid32.as<float>(), t->N, V[p], nv[p], idx[p], nidx[p], kind[p], lambda[p], nullptr, V[p]};
  }
  PhaseTimer timer;
  hipStream_t s = as_stream(stream);
  int rc;
  if ((rc = timer.start(o.timing != 0, s))) return rc;
  if ((rc = timer.mark(1))) return rc;
  if ((rc = lm_core(pairs, o, stats, s, timer, ext, ext_cap))) return rc;
  return timer.finish();
}

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];           // one call at a time per device (they share its array)
char* g_scratch_ptr[kMaxDevices] = {nullptr};   // set under g_scratch_mu[device]

}  // namespace

// Scratch of mo_deform_batch_noalloc: one array per GPU, held until the
// process ends.  Host code never passes `scratch` itself (in host code the
// name is a stand-in, not a device address): each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

MO_EXPORT int mo_deform_batch(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                              const int32_t* const* idx, const int64_t* nidx, const int32_t* kind,
                              const double* lambda, const mo_lm_options* opt, mo_lm_stats* stats, void* stream) {
  return deform_batch_impl(npairs, param_ids, V, nv, idx, nidx, kind, lambda, opt, stats, stream, nullptr, 0);
}

MO_EXPORT int64_t mo_deform_batch_noalloc_max_bytes(void) { return (int64_t)kNoAllocScratchBytes; }

MO_EXPORT int mo_deform_batch_noalloc(int32_t npairs, const int32_t* param_ids, double* const* V,
                                      const int64_t* nv, const int32_t* const* idx, const int64_t* nidx,
                                      const int32_t* kind, const double* lambda, const mo_lm_options* opt,
                                      mo_lm_stats* stats, void* stream) {
  int dev = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_CHECK(dev >= 0 && dev < kMaxDevices, MO_ERR_ARG, "mo_deform_batch_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MO_CHECK(p != nullptr, MO_ERR_HIP, "mo_deform_batch_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (char*)p;
  }
  const int rc = deform_batch_impl(npairs, param_ids, V, nv, idx, nidx, kind, lambda, opt, stats, stream,
                                   g_scratch_ptr[dev], kNoAllocScratchBytes);
  // the array is reused by the next call: nothing of this call may still run
  (void)hipStreamSynchronize(as_stream(stream));
  return rc;
}

namespace {

int rigid_deform_batch_impl(int32_t npairs, const double* const* src, const int64_t* src_nv,
                            const int32_t* const* src_faces, const int64_t* src_nf, const double* const* tgt,
                            const int64_t* tgt_nv, const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                            const int32_t* grid_resolution, const double* lambda, const mo_lm_options& o,
                            double* const* out, mo_lm_stats* stats, hipStream_t s) {
  // per-pair scratch: normalized target, (scale, pos), normalized source, FP32 grid
  std::vector<GridSpec> gspecs(npairs);
  size_t bytes = 512;
  for (int p = 0; p < npairs; ++p) {
    MO_CHECK(tgt[p] && tgt_faces[p] && tgt_nv[p] > 0 && tgt_nf[p] > 0, MO_ERR_ARG, "pair %d: empty target", p);
    MO_CHECK(src_nv[p] >= 0 && (src[p] || src_nv[p] == 0), MO_ERR_ARG, "pair %d: bad source", p);
    MO_CHECK(grid_resolution[p] >= 2 && grid_resolution[p] <= 1024, MO_ERR_ARG, "pair %d: bad grid resolution", p);
    long long N = grid_resolution[p];
    bytes += al(tgt_nv[p] * 3 * sizeof(double)) + al(8 * sizeof(double)) +
             al(std::max<long long>(src_nv[p], 1) * 3 * sizeof(double)) + al(N * N * N * sizeof(float));
    gspecs[p] = GridSpec{nullptr, tgt_nv[p], tgt_faces[p], tgt_nf[p], (int)N, nullptr, nullptr, nullptr,
                         lookup_band2((int)N)};
  }
  const size_t gws = grid_batch_workspace_bytes(gspecs.data(), npairs);
  // grid work counters: one 64-byte line per pair (atomics from that
  // pair's wavefronts only), summed on the host
  constexpr int kStatStride = 8;
  bytes += al(gws) + 2 * al(xform_workspace_bytes(npairs)) +
           al((size_t)npairs * kStatStride * sizeof(unsigned long long));
  Arena ar;
  int rc;
  PhaseTimer timer;
  if ((rc = timer.start(o.timing != 0, s))) return rc;
  if ((rc = ar.reserve(bytes, s))) return rc;
  int* err = ar.take<int>(1);
  // with timing: the grid build's work counters (summed over the batch)
  unsigned long long* gstats = ar.take<unsigned long long>((size_t)npairs * kStatStride);
  if (o.timing == 2) {
    if ((rc = zero_fill(gstats, (size_t)npairs * kStatStride * sizeof(unsigned long long), s))) return rc;
    for (int p = 0; p < npairs; ++p) gspecs[p].stats = gstats + (size_t)p * kStatStride;
  }
  void* gws_ptr = ar.take<char>(gws);
  void* xws_t = ar.take<char>(xform_workspace_bytes(npairs));
  void* xws_s = ar.take<char>(xform_workspace_bytes(npairs));
  if ((rc = zero_fill(err, sizeof(int), s))) return rc;
  std::vector<CorePair> pairs(npairs);
  std::vector<XformJob> tj(npairs), sj(npairs);
  for (int p = 0; p < npairs; ++p) {
    const long long N = grid_resolution[p];
    double* tn = ar.take<double>(tgt_nv[p] * 3);
    double* sp = ar.take<double>(8);
    double* sn = ar.take<double>(std::max<long long>(src_nv[p], 1) * 3);
    float* g32 = ar.take<float>(N * N * N);
    tj[p] = XformJob{(const void*)tgt[p], 1, tgt_nv[p], sp, tn};
    sj[p] = XformJob{(const void*)src[p], 1, src_nv[p], sp, sn};
    gspecs[p].Vn = tn;
    gspecs[p].dist32 = g32;
    pairs[p] = CorePair{g32, (int)N, sn, src_nv[p], src_faces[p], src_nf[p], 0, lambda[p], sp, out[p]};
  }
  // normalization and the sources' topology first; the grid builds are
  // enqueued while the host waits for the topology's sizes (they steer the
  // solver launches, which then queue behind the grids)
  if ((rc = xform_batch_launch(tj.data(), npairs, true, xws_t, s))) return rc;
  if ((rc = xform_batch_launch(sj.data(), npairs, false, xws_s, s))) return rc;
  timer.setup_first = true;
  auto grids = [&]() -> int {
    int r = timer.mark(1);
    if (r) return r;
    // after another batch's solver launches (asynchronous batches): the
    // face preparation runs at once, the grid build proper waits
    return grid_build_batch_launch(gspecs.data(), npairs, gws_ptr, s, err, g_async ? g_async->after : nullptr);
  };
  if ((rc = lm_core(pairs, o, stats, s, timer, nullptr, 0, grids))) return rc;
  if ((rc = timer.finish())) return rc;
  if (o.timing == 2) {
    std::vector<unsigned long long> h((size_t)npairs * kStatStride);
    MO_HIP(hipMemcpyAsync(h.data(), gstats, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MO_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < kGridStatCount; ++k) {
      g_grid_stats[k] = 0;
      for (int p = 0; p < npairs; ++p) g_grid_stats[k] += (long long)h[(size_t)p * kStatStride + k];
    }
  }
  if (AsyncCtx* a = g_async) {
    if (a->deferred) {
      MO_HIP(hipMemcpyAsync(a->pinned + a->err_off, err, sizeof(int), hipMemcpyDeviceToHost, s));
      return MO_OK;
    }
  }
  return check_err_flag(err, s, "target faces");
}

int check_rigid_args(int32_t npairs, const double* const* src, const int64_t* src_nv, const int32_t* const* src_faces,
                     const int64_t* src_nf, const double* const* tgt, const int64_t* tgt_nv,
                     const int32_t* const* tgt_faces, const int64_t* tgt_nf, const int32_t* grid_resolution,
                     const double* lambda, double* const* out) {
  MO_CHECK(npairs >= 0, MO_ERR_ARG, "npairs < 0");
  if (npairs == 0) return MO_OK;
  MO_CHECK(src && src_nv && src_faces && src_nf && tgt && tgt_nv && tgt_faces && tgt_nf && grid_resolution &&
               lambda && out,
           MO_ERR_ARG, "null per-pair array");
  return MO_OK;
}

}  // namespace

MO_EXPORT int mo_rigid_deform_batch(int32_t npairs, const double* const* src, const int64_t* src_nv,
                                    const int32_t* const* src_faces, const int64_t* src_nf, const double* const* tgt,
                                    const int64_t* tgt_nv, const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                                    const int32_t* grid_resolution, const double* lambda, const mo_lm_options* opt,
                                    double* const* out, mo_lm_stats* stats, void* stream) {
  int rc = check_rigid_args(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf, grid_resolution,
                            lambda, out);
  if (rc || npairs == 0) return rc;
  mo_lm_options o;
  if (opt) o = *opt; else mo_lm_default_options(&o);
  HostStage stage(as_stream(stream));
  StageScope scope(&stage);
  rc = rigid_deform_batch_impl(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf,
                               grid_resolution, lambda, o, out, stats, as_stream(stream));
  stage.drained = rc == MO_OK;   // a successful call ends with a synchronization
  return rc;
}

// ---------------------------------------------------------------------------
// Batch graphs (see ChainCtx).

struct mo_batch_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  char* pool = nullptr;
  char* pinned = nullptr;
  mo::ChainCtx ctx;
  ~mo_batch_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pool) (void)hipFree(pool);
    if (pinned) (void)hipHostFree(pinned);
  }
};

namespace {

struct ChainScope {
  explicit ChainScope(ChainCtx* c) { g_chain = c; }
  ~ChainScope() { g_chain = nullptr; }
};

int batch_graph_finish(mo_batch_graph* g, mo_lm_stats* stats) {
  const ChainCtx& c = g->ctx;
  for (const auto& ck : c.checks) {
    const std::vector<char>& want = c.reads[ck.second];
    MO_CHECK(memcmp(g->pinned + ck.first, want.data(), want.size()) == 0, MO_ERR_STATE,
             "batch graph: the batch's topology, faces or an error flag differ from the captured run "
             "(results invalid; the inputs' faces must not change between replays)");
  }
  if (stats)
    fill_stats(reinterpret_cast<const LMResultDev*>(g->pinned + c.res_off), c.res_path.data(), c.res_wgs.data(),
               c.res_np, stats);
  return MO_OK;
}

}  // namespace

MO_EXPORT int mo_rigid_deform_batch_graph(int32_t npairs, const double* const* src, const int64_t* src_nv,
                                          const int32_t* const* src_faces, const int64_t* src_nf,
                                          const double* const* tgt, const int64_t* tgt_nv,
                                          const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                                          const int32_t* grid_resolution, const double* lambda,
                                          const mo_lm_options* opt, double* const* out, mo_lm_stats* stats,
                                          void* stream, mo_batch_graph** graph) {
  MO_CHECK(graph, MO_ERR_ARG, "graph must not be null");
  *graph = nullptr;
  int rc = check_rigid_args(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf, grid_resolution,
                            lambda, out);
  if (rc) return rc;
  MO_CHECK(npairs > 0, MO_ERR_ARG, "a batch graph needs at least one pair");
  for (int p = 0; p < npairs; ++p)
    MO_CHECK((const void*)out[p] != (const void*)src[p], MO_ERR_ARG,
             "pair %d: out aliases src (a replay reads the sources again)", p);
  mo_lm_options o;
  if (opt) o = *opt; else mo_lm_default_options(&o);
  MO_CHECK(o.timing == 0, MO_ERR_ARG, "options.timing must be 0 for a batch graph");
  std::unique_ptr<mo_batch_graph> g(new mo_batch_graph);
  MO_HIP(hipGetDevice(&g->device));
  hipStream_t s = as_stream(stream);
  // 1. the eager run (its results and stats are the call's), recording the
  //    steering values and sizes
  g->ctx.mode = ChainCtx::kRecord;
  {
    ChainScope scope(&g->ctx);
    if ((rc = rigid_deform_batch_impl(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf,
                                      grid_resolution, lambda, o, out, stats, s)))
      return rc;
  }
  MO_CHECK(!g->ctx.team_launch, MO_ERR_ARG,
           "batch graph: a pair of this batch needs a cooperative multi-workgroup launch, which is not captured "
           "(the eager run's results and stats are valid; use mo_rigid_deform_batch for such batches)");
  // 2. the graph's own memory, then the capture on a private stream
  ChainCtx& c = g->ctx;
  c.pool_cap = c.pool_bytes;
  c.pinned_cap = c.pinned_bytes;
  MO_HIP(hipMalloc((void**)&g->pool, std::max<size_t>(c.pool_cap, 256)));
  MO_HIP(hipHostMalloc((void**)&g->pinned, std::max<size_t>(c.pinned_cap, 256), hipHostMallocDefault));
  c.pool = g->pool;
  c.pinned = g->pinned;
  c.mode = ChainCtx::kCapture;
  hipStream_t cs = nullptr;
  MO_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{cs};
  MO_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  {
    ChainScope scope(&c);
    rc = rigid_deform_batch_impl(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf,
                                 grid_resolution, lambda, o, out, nullptr, cs);
  }
  hipGraph_t captured = nullptr;
  const hipError_t ec = hipStreamEndCapture(cs, &captured);
  if (rc) {
    if (captured) (void)hipGraphDestroy(captured);
    return rc;
  }
  MO_CHECK(ec == hipSuccess && captured, MO_ERR_HIP, "batch graph: stream capture failed: %s", hipGetErrorString(ec));
  g->graph = captured;
  MO_CHECK(c.next_read == c.reads.size(), MO_ERR_STATE, "batch graph capture diverged from the recorded run");
  MO_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *graph = g.release();
  return MO_OK;
}

MO_EXPORT int mo_batch_graph_launch(mo_batch_graph* graph, mo_lm_stats* stats, void* stream) {
  MO_CHECK(graph && graph->exec, MO_ERR_ARG, "null batch graph");
  int dev = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_CHECK(dev == graph->device, MO_ERR_STATE, "batch graph was captured on device %d, current device is %d",
           graph->device, dev);
  hipStream_t s = as_stream(stream);
  MO_HIP(hipGraphLaunch(graph->exec, s));
  MO_HIP(hipStreamSynchronize(s));
  return batch_graph_finish(graph, stats);
}

MO_EXPORT int mo_batch_graph_destroy(mo_batch_graph* graph) {
  if (!graph) return MO_OK;
  int dev = 0;
  MO_HIP(hipGetDevice(&dev));
  MO_CHECK(dev == graph->device, MO_ERR_STATE, "batch graph belongs to device %d, current device is %d",
           graph->device, dev);
  delete graph;
  return MO_OK;
}

// ---------------------------------------------------------------------------
// Asynchronous batches (see AsyncCtx).

namespace {

struct AsyncScope {
  explicit AsyncScope(AsyncCtx* a) { g_async = a; }
  ~AsyncScope() { g_async = nullptr; }
};

}  // namespace

struct mo_batch_handle {
  hipEvent_t done = nullptr;
  char* pinned = nullptr;
  size_t pinned_cap = 0;
  mo::AsyncCtx ctx;
  std::vector<mo_lm_stats> stats;   // the batch completed inside the call (team launches)
  int npairs = 0;
  std::unique_ptr<mo::HostStage> stage;
  ~mo_batch_handle() {
    stage.reset();   // waits for the stream unless the batch is known to be complete
    if (done) (void)hipEventDestroy(done);
    pinned_put(pinned, pinned_cap);
  }
};

MO_EXPORT int mo_rigid_deform_batch_async(int32_t npairs, const double* const* src, const int64_t* src_nv,
                                          const int32_t* const* src_faces, const int64_t* src_nf,
                                          const double* const* tgt, const int64_t* tgt_nv,
                                          const int32_t* const* tgt_faces, const int64_t* tgt_nf,
                                          const int32_t* grid_resolution, const double* lambda,
                                          const mo_lm_options* opt, double* const* out, void* stream,
                                          mo_batch_handle* after, mo_batch_handle** handle) {
  MO_CHECK(handle, MO_ERR_ARG, "handle must not be null");
  *handle = nullptr;
  int rc = check_rigid_args(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf, grid_resolution,
                            lambda, out);
  if (rc) return rc;
  mo_lm_options o;
  if (opt) o = *opt; else mo_lm_default_options(&o);
  MO_CHECK(o.timing == 0, MO_ERR_ARG, "options.timing must be 0 for an asynchronous batch");
  std::unique_ptr<mo_batch_handle> h(new mo_batch_handle);
  const size_t res_bytes = al256c((size_t)std::max(npairs, 1) * sizeof(LMResultDev));
  if ((rc = pinned_get(res_bytes + 256, &h->pinned, &h->pinned_cap))) return rc;
  h->ctx.pinned = h->pinned;
  h->ctx.err_off = res_bytes;
  h->stats.resize((size_t)std::max(npairs, 1));
  h->npairs = npairs;
  hipStream_t s = as_stream(stream);
  h->stage.reset(new HostStage(s));
  if (after) h->ctx.after = after->ctx.lm_ready;   // null: that batch launched no solver
  {
    AsyncScope scope(&h->ctx);
    StageScope stage_scope(h->stage.get());
    if ((rc = rigid_deform_batch_impl(npairs, src, src_nv, src_faces, src_nf, tgt, tgt_nv, tgt_faces, tgt_nf,
                                      grid_resolution, lambda, o, out, h->stats.data(), s)))
      return rc;
  }
  if (!h->ctx.deferred) {
    h->stage->drained = true;   // completed inside the call
  } else {
    MO_HIP(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
    MO_HIP(hipEventRecord(h->done, s));
  }
  h->ctx.after = nullptr;
  *handle = h.release();
  return MO_OK;
}

MO_EXPORT int mo_batch_wait(mo_batch_handle* handle, mo_lm_stats* stats) {
  MO_CHECK(handle, MO_ERR_ARG, "null batch handle");
  std::unique_ptr<mo_batch_handle> h(handle);
  const AsyncCtx& a = h->ctx;
  if (!a.deferred) {
    if (stats) std::copy(h->stats.begin(), h->stats.begin() + h->npairs, stats);
    return MO_OK;
  }
  MO_HIP(hipEventSynchronize(h->done));
  h->stage->drained = true;   // every staged copy precedes the event
  int herr = 0;
  memcpy(&herr, h->pinned + a.err_off, sizeof(int));
  MO_CHECK(herr == 0, MO_ERR_RANGE, "target faces: index out of range (flags 0x%x)", herr);
  if (stats) fill_stats(reinterpret_cast<const LMResultDev*>(h->pinned), a.path.data(), a.wgs.data(), a.np, stats);
  return MO_OK;
}

namespace {

// DeformSubdivision's residual order: the geometry-neighbour pairs, then
// every face half-edge (deformer.cc:194-226).
__global__ void k_subdivision_edges(const int32_t* __restrict__ E, long long ne, const int32_t* __restrict__ F,
                                    long long nf, int32_t* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) {
    out[2 * i] = E[2 * i];
    out[2 * i + 1] = E[2 * i + 1];
  } else if (i < ne + 3 * nf) {
    const long long h = i - ne, f = h / 3;
    const int j = (int)(h - 3 * f);
    out[2 * i] = F[3 * f + j];
    out[2 * i + 1] = F[3 * f + (j + 1) % 3];
  }
}

}  // namespace

MO_EXPORT int mo_deform_subdivision(int32_t param_id, const double* V_rest, double* V, int64_t nv, const int32_t* F,
                                    int64_t nf, const int32_t* E, int64_t ne, double lambda,
                                    const mo_lm_options* opt, int32_t max_solves, mo_flow_callback callback,
                                    void* user, mo_lm_stats* stats, int32_t* solves, void* stream) {
  Template* t = get_template(param_id);
  MO_CHECK(t, MO_ERR_ARG, "unknown param_id %d", param_id);
  MO_CHECK(nv >= 0 && nf >= 0 && ne >= 0 && max_solves > 0, MO_ERR_ARG, "mo_deform_subdivision: bad sizes");
  MO_CHECK((V_rest && V) || nv == 0, MO_ERR_ARG, "mo_deform_subdivision: null vertex buffer");
  MO_CHECK((F || nf == 0) && (E || ne == 0), MO_ERR_ARG, "mo_deform_subdivision: null index buffer");
  if (solves) *solves = 0;
  if (nv == 0) return MO_OK;
  mo_lm_options o;
  if (opt) o = *opt; else mo_lm_default_options(&o);
  hipStream_t s = as_stream(stream);
  const long long nh = ne + 3 * nf;
  Arena ar;
  int rc;
  if ((rc = ar.reserve(al(std::max<long long>(nh, 1) * 2 * sizeof(int32_t)) + al(nv * 3 * sizeof(double)) + 512, s)))
    return rc;
  int32_t* edges = ar.take<int32_t>(std::max<long long>(nh, 1) * 2);
  double* start = ar.take<double>(nv * 3);
  if (nh > 0) {
    hipLaunchKernelGGL(k_subdivision_edges, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, E, (long long)ne, F,
                       (long long)nf, edges);
    MO_HIP(hipGetLastError());
  }
  // TerminateWhenSuccessCallback keeps its counter across solves: a solve
  // that starts with the counter at 0 calls back at its first successful
  // step and stops at its second; one that starts at 1 stops at its first.
  int counter = 0;
  double prev = 1e30;
  mo_lm_stats st{};
  int n = 0;
  for (; n < max_solves;) {
    MO_HIP(hipMemcpyAsync(start, V, nv * 3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    CorePair c{t->grid32.as<float>(), t->N, V_rest, nv, edges, nh, 2, lambda, nullptr, V, start};
    mo_lm_options oi = o;
    oi.max_successful_steps = counter == 0 ? 2 : 1;
    PhaseTimer timer;
    if ((rc = lm_core(std::vector<CorePair>{c}, oi, &st, s, timer))) return rc;
    ++n;
    const int k = st.successful_steps;
    if (counter == 0) {
      if (k >= 1 && callback) callback(start, nv, user);
      counter = k == 1 ? 1 : 0;
    } else {
      counter = k >= 1 ? 0 : 1;
    }
    if (std::fabs(prev - st.final_cost) < 1e-6) break;
    prev = st.final_cost;
  }
  if (stats) *stats = st;
  if (solves) *solves = n;
  return MO_OK;
}

// ---------------------------------------------------------------------------
// Deformer::DeformWithRot (deformer.cc:94-171).

namespace {

__global__ void k_rot_init(const double* __restrict__ V, const double* __restrict__ rot, int nv,
                           float4* __restrict__ x0, float4* __restrict__ u, float4* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  x0[i] = make_float4((float)V[3 * i], (float)V[3 * i + 1], (float)V[3 * i + 2], 0.f);
  u[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  w[i] = rot ? make_float4((float)rot[3 * i], (float)rot[3 * i + 1], (float)rot[3 * i + 2], 0.f)
             : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_rot_finish(double* __restrict__ V, double* __restrict__ rot, int nv, const float4* __restrict__ u,
                             const float4* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const float4 d = u[i];
  V[3 * i] += (double)d.x;
  V[3 * i + 1] += (double)d.y;
  V[3 * i + 2] += (double)d.z;
  if (rot) {
    const float4 r = w[i];
    rot[3 * i] = r.x;
    rot[3 * i + 1] = r.y;
    rot[3 * i + 2] = r.z;
  }
}

}  // namespace

MO_EXPORT int mo_deform_rot_batch(int32_t npairs, const int32_t* param_ids, double* const* V, const int64_t* nv,
                                  const int32_t* const* F, const int64_t* nf, const double* lambda,
                                  const mo_lm_options* opt, double* const* rot, mo_lm_stats* stats, void* stream) {
  MO_CHECK(npairs >= 0, MO_ERR_ARG, "mo_deform_rot_batch: npairs < 0");
  if (npairs == 0) return MO_OK;
  MO_CHECK(param_ids && V && nv && F && nf && lambda, MO_ERR_ARG, "mo_deform_rot_batch: null per-pair array");
  mo_lm_options o;
  if (opt) {
    o = *opt;
  } else {
    // the rotation unknowns make the path sensitive to inexact steps: a
    // tighter inner tolerance than Deform's keeps it on the exact-step path
    mo_lm_default_options(&o);
    o.pcg_relative_tolerance = 1e-6f;
  }
  MO_CHECK(o.block_threads >= 64 && o.block_threads <= 1024 && o.block_threads % 64 == 0, MO_ERR_ARG,
           "block_threads must be a multiple of 64 in [64, 1024]");
  hipStream_t s = as_stream(stream);
  std::vector<RotPairDesc> pd(npairs);
  std::vector<std::vector<int32_t>> hF(npairs);
  for (int p = 0; p < npairs; ++p) {
    Template* t = get_template(param_ids[p]);
    MO_CHECK(t, MO_ERR_ARG, "mo_deform_rot: pair %d: unknown param_id %d", p, param_ids[p]);
    MO_CHECK(V[p] && nv[p] > 0 && nv[p] < (1LL << 24) && nf[p] >= 0 && nf[p] < (1LL << 28) && (F[p] || nf[p] == 0),
             MO_ERR_ARG, "mo_deform_rot: pair %d: bad arguments (nv %lld, nf %lld)", p, (long long)nv[p],
             (long long)nf[p]);
    pd[p].grid = t->grid32.as<float>();
    pd[p].N = t->N;
    pd[p].lambda = (float)lambda[p];
    hF[p].resize((size_t)nf[p] * 3);
    if (nf[p]) MO_HIP(hipMemcpyAsync(hF[p].data(), F[p], hF[p].size() * 4, hipMemcpyDeviceToHost, s));
  }
  MO_HIP(hipStreamSynchronize(s));
  // per pair: half-edges in residual order (face, j), skipping a == b (Ceres
  // rejects a residual block that names one parameter block twice), and
  // each vertex's outgoing then incoming half-edges
  std::vector<int2> hab;
  std::vector<int4> vadj;
  std::vector<int> vlist;
  long long vtot = 0, htot = 0;
  for (int p = 0; p < npairs; ++p) {
    const int n = (int)nv[p];
    for (size_t k = 0; k < hF[p].size(); ++k)
      MO_CHECK(hF[p][k] >= 0 && hF[p][k] < n, MO_ERR_ARG, "mo_deform_rot: pair %d: face index %d out of range", p,
               hF[p][k]);
    const size_t h0 = hab.size();
    for (int64_t f = 0; f < nf[p]; ++f)
      for (int j = 0; j < 3; ++j) {
        const int a = hF[p][3 * f + j], b = hF[p][3 * f + (j + 1) % 3];
        if (a != b) hab.push_back(make_int2(a, b));
      }
    const int nh = (int)(hab.size() - h0);
    std::vector<int> nout(n, 0), nin(n, 0);
    for (size_t h = h0; h < hab.size(); ++h) { nout[hab[h].x]++; nin[hab[h].y]++; }
    const size_t v0 = vadj.size(), l0 = vlist.size();
    int pos = 0;
    for (int i = 0; i < n; ++i) {
      vadj.push_back(make_int4(pos, nout[i], nin[i], 0));
      pos += nout[i] + nin[i];
    }
    vlist.resize(l0 + (size_t)pos);
    std::vector<int> fo(n), fi(n);
    for (int i = 0; i < n; ++i) { fo[i] = vadj[v0 + i].x; fi[i] = vadj[v0 + i].x + vadj[v0 + i].y; }
    for (int h = 0; h < nh; ++h) {
      vlist[l0 + fo[hab[h0 + h].x]++] = h;
      vlist[l0 + fi[hab[h0 + h].y]++] = h;
    }
    pd[p].nv = n;
    pd[p].nh = nh;
    pd[p].vbase = vtot;
    pd[p].hbase = htot;
    pd[p].lbase = (long long)l0;
    vtot += n;
    htot += nh;
  }
  const long long hslots = htot ? htot : 1;
  DevPtr d_x0, d_u, d_w, d_adj, d_list, d_hab, d_vw, d_hw, d_pair, d_res;
  if (int e_ = d_x0.alloc((size_t)vtot * 16)) return e_;
  if (int e_ = d_u.alloc((size_t)vtot * 16)) return e_;
  if (int e_ = d_w.alloc((size_t)vtot * 16)) return e_;
  if (int e_ = d_adj.alloc((size_t)vtot * 16)) return e_;
  if (int e_ = d_list.alloc(vlist.size() * 4)) return e_;
  if (int e_ = d_hab.alloc((size_t)htot * 8)) return e_;
  if (int e_ = d_vw.alloc((size_t)vtot * 16 * kRotVSlots)) return e_;
  if (int e_ = d_hw.alloc((size_t)hslots * 16 * kRotHSlots)) return e_;
  if (int e_ = d_pair.alloc(npairs * sizeof(RotPairDesc))) return e_;
  if (int e_ = d_res.alloc(npairs * sizeof(LMResultDev))) return e_;
  MO_HIP(hipMemcpyAsync(d_adj.p, vadj.data(), (size_t)vtot * 16, hipMemcpyHostToDevice, s));
  if (!vlist.empty()) MO_HIP(hipMemcpyAsync(d_list.p, vlist.data(), vlist.size() * 4, hipMemcpyHostToDevice, s));
  if (htot) MO_HIP(hipMemcpyAsync(d_hab.p, hab.data(), (size_t)htot * 8, hipMemcpyHostToDevice, s));
  MO_HIP(hipMemcpyAsync(d_pair.p, pd.data(), npairs * sizeof(RotPairDesc), hipMemcpyHostToDevice, s));
  for (int p = 0; p < npairs; ++p) {
    const unsigned blocks = (unsigned)((nv[p] + 255) / 256);
    hipLaunchKernelGGL(k_rot_init, dim3(blocks), dim3(256), 0, s, V[p], rot ? rot[p] : nullptr, (int)nv[p],
                       d_x0.as<float4>() + pd[p].vbase, d_u.as<float4>() + pd[p].vbase,
                       d_w.as<float4>() + pd[p].vbase);
  }
  MO_HIP(hipGetLastError());
  RotArrays A{d_x0.as<float4>(), d_u.as<float4>(), d_w.as<float4>(), d_adj.as<int4>(), d_list.as<int>(),
              d_hab.as<int2>(), d_vw.as<float4>(), d_hw.as<float4>()};
  // team_vertices > 0: each problem on a team of ceil(nv / team_vertices)
  // co-resident workgroups (k_lm_rot_team; at most lm_team_max()), packed
  // largest first into cooperative launches, one such launch at a time in
  // the process (team_launch_mutex); otherwise one workgroup per problem.
  const int rbt = std::min(o.block_threads, 256);
  std::unique_lock<std::mutex> team_lock(team_launch_mutex(), std::defer_lock);
  struct TeamDrain {
    std::unique_lock<std::mutex>& lk;
    hipStream_t s;
    ~TeamDrain() {
      if (lk.owns_lock()) (void)hipStreamSynchronize(s);
    }
  } team_drain{team_lock, s};
  int path = MO_PATH_ONE_WORKGROUP;
  std::vector<int> rot_wgs(npairs, 1);   // workgroups that solve each problem
  DevPtr d_team;
  int* d_abort = nullptr;
  int rc;
  MO_CHECK(o.team_vertices >= 0, MO_ERR_ARG, "team_vertices must be >= 0");
  const int rcap = o.team_vertices > 0 ? lm_rot_team_capacity(rbt) : 0;
  if (o.team_vertices > 0 && rcap >= lm_team_max()) {
    std::vector<int> order(npairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nv[a] > nv[b]; });
    std::vector<int> tc(npairs);
    long long nwg_all = 0;
    for (int t = 0; t < npairs; ++t) {
      tc[t] = (int)std::min<long long>(std::max<long long>((nv[order[t]] + o.team_vertices - 1) / o.team_vertices, 1),
                                       lm_team_max());
      nwg_all += tc[t];
      rot_wgs[order[t]] = tc[t];
    }
    std::vector<int4> wg;
    std::vector<int> chunk_team{0}, chunk_wg{0};
    int used = 0;
    for (int t = 0; t < npairs; ++t) {
      if (used + tc[t] > rcap) {
        chunk_team.push_back(t);
        chunk_wg.push_back((int)wg.size());
        used = 0;
      }
      for (int r = 0; r < tc[t]; ++r) wg.push_back(make_int4(t - chunk_team.back(), r, tc[t], 0));
      used += tc[t];
    }
    chunk_team.push_back(npairs);
    chunk_wg.push_back((int)wg.size());
    const int per = lm_team_part_doubles();
    const size_t b_tp = al(npairs * sizeof(int)), b_wg = al(nwg_all * sizeof(int4)),
                 b_cnt = al(((size_t)npairs * 32 + 1) * sizeof(unsigned)), b_part = al((size_t)npairs * per * 8);
    if (int e_ = d_team.alloc(b_tp + b_wg + b_cnt + b_part)) return e_;
    char* base = d_team.as<char>();
    int* d_tp = (int*)base;
    int4* d_wg = (int4*)(base + b_tp);
    unsigned* d_cnt = (unsigned*)(base + b_tp + b_wg);
    double* d_part = (double*)(base + b_tp + b_wg + b_cnt);
    d_abort = (int*)(d_cnt + (size_t)npairs * 32);
    if ((rc = zero_fill(d_cnt, ((size_t)npairs * 32 + 1) * sizeof(unsigned), s))) return rc;
    MO_HIP(hipMemcpyAsync(d_tp, order.data(), npairs * sizeof(int), hipMemcpyHostToDevice, s));
    MO_HIP(hipMemcpyAsync(d_wg, wg.data(), wg.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    MO_HIP(hipStreamSynchronize(s));   // host vectors above go out of scope
    team_lock.lock();
    path = MO_PATH_TEAM;
    for (size_t k = 0; k + 1 < chunk_team.size(); ++k) {
      const int t0 = chunk_team[k];
      rc = lm_rot_team_launch(d_pair.as<RotPairDesc>(), d_tp + t0, d_wg + chunk_wg[k], chunk_wg[k + 1] - chunk_wg[k],
                              A, to_dev_options(o), d_res.as<LMResultDev>(), vtot, hslots, rbt,
                              d_cnt + (size_t)t0 * 32, d_part + (size_t)t0 * per, d_abort, s);
      if (rc == kCoopRefused && k == 0) {
        // the runtime refused the cooperative grid before any problem ran:
        // one workgroup per problem instead
        if ((rc = lm_rot_launch(d_pair.as<RotPairDesc>(), npairs, A, to_dev_options(o), d_res.as<LMResultDev>(), vtot,
                                hslots, rbt, s)))
          return rc;
        path = MO_PATH_COOP_REFUSED;
        std::fill(rot_wgs.begin(), rot_wgs.end(), 1);
        break;
      }
      if (rc == kCoopRefused) return fail(MO_ERR_HIP, "DeformWithRot: cooperative launch refused after %zu chunks", k);
      if (rc) return rc;
    }
  } else {
    rc = lm_rot_launch(d_pair.as<RotPairDesc>(), npairs, A, to_dev_options(o), d_res.as<LMResultDev>(), vtot, hslots,
                       rbt, s);
    if (rc) return rc;
  }
  if (d_abort) {
    // invalid solutions after a barrier wait ran out: raise before the
    // caller's V / rot are overwritten
    int h_abort = 0;
    MO_HIP(hipMemcpyAsync(&h_abort, d_abort, sizeof(int), hipMemcpyDeviceToHost, s));
    MO_HIP(hipStreamSynchronize(s));
    MO_CHECK(h_abort == 0, MO_ERR_HIP,
             "DeformWithRot team solve: a barrier wait exceeded its bound (results invalid; outputs not written)");
  }
  for (int p = 0; p < npairs; ++p) {
    const unsigned blocks = (unsigned)((nv[p] + 255) / 256);
    hipLaunchKernelGGL(k_rot_finish, dim3(blocks), dim3(256), 0, s, V[p], rot ? rot[p] : nullptr, (int)nv[p],
                       d_u.as<float4>() + pd[p].vbase, d_w.as<float4>() + pd[p].vbase);
  }
  MO_HIP(hipGetLastError());
  std::vector<LMResultDev> r(npairs);
  MO_HIP(hipMemcpyAsync(r.data(), d_res.p, npairs * sizeof(LMResultDev), hipMemcpyDeviceToHost, s));
  MO_HIP(hipStreamSynchronize(s));
  if (stats)
    for (int p = 0; p < npairs; ++p) {
      mo_lm_stats& o2 = stats[p];
      o2.initial_cost = r[p].initial_cost;
      o2.final_cost = r[p].final_cost;
      o2.vertex_cost = r[p].vertex_cost;
      o2.rigidity_cost = r[p].rigidity_cost;
      o2.iterations = r[p].iterations;
      o2.successful_steps = r[p].successful_steps;
      o2.unsuccessful_steps = r[p].unsuccessful_steps;
      o2.termination = r[p].termination;
      o2.pcg_iterations = r[p].pcg_iterations;
      o2.clock_start = r[p].clock_start;
      o2.clock_end = r[p].clock_end;
      o2.solver_path = path;
      o2.workgroups = rot_wgs[p];
    }
  return MO_OK;
}

MO_EXPORT int mo_deform_rot(int32_t param_id, double* V, int64_t nv, const int32_t* F, int64_t nf, double lambda,
                            const mo_lm_options* opt, double* rot, mo_lm_stats* stats, void* stream) {
  double* Vp[1] = {V};
  double* Rp[1] = {rot};
  const int32_t* Fp[1] = {F};
  return mo_deform_rot_batch(1, &param_id, Vp, &nv, Fp, &nf, &lambda, opt, rot ? Rp : nullptr, stats, stream);
}
