// Shared host/device definitions for the MeshODE MI355X library.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/meshode.h"

#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>

namespace mo {

// ---------------------------------------------------------------------------
// Error reporting: every C-ABI entry point returns an int status (MO_OK or a
// negative MO_ERR_* code from meshode.h) and records a message retrievable by
// mo_last_error().
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);

#define MO_HIP(call)                                                              \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      return ::mo::fail(MO_ERR_HIP, "%s failed: %s (%s:%d)", #call,         \
                        hipGetErrorString(_e), __FILE__, __LINE__);               \
  } while (0)

#define MO_CHECK(cond, code, ...)                                                 \
  do {                                                                            \
    if (!(cond)) return ::mo::fail((code), __VA_ARGS__);                          \
  } while (0)

// ---------------------------------------------------------------------------
// Host->device copies of a batch's launch chain.  Outside graph capture this
// is hipMemcpyAsync; while a batch graph is being captured (mo_capi.hip) the
// bytes are first copied into pinned memory the graph owns, so that every
// replay copies the same bytes.
int chain_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);

// ---------------------------------------------------------------------------
// Float -> int conversion with x86 cvtt semantics (out of int32 range or NaN
// -> INT_MIN).  The reference computes `int px = p * N` (uniformgrid.cc:20-22,
// 87-89); on its x86 build an out-of-range value yields INT_MIN.
__host__ __device__ inline int trunc_i32f(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
  return (int)f;
}
__host__ __device__ inline int trunc_i32d(double f) {
  if (!(f > -2147483649.0 && f < 2147483648.0)) return INT_MIN;
  return (int)f;
}

// Per-template grid build summary kept on the host.
struct GridDims {
  int N;
  long long cells;
};

}  // namespace mo
