// Device helpers shared by the LM solver kernels: block-wide FP64
// reductions and the trust-region loop itself (Ceres' TrustRegionMinimizer
// with the Levenberg-Marquardt strategy, restated; see mo_lm.hip).
#pragma once

#include "mo_common.h"
#include "mo_lm.h"

namespace mo {

constexpr int kMaxWaves = 16;

// Block-wide reduction of K sums followed by M maxima; every thread gets the
// totals, summed in the same order by every thread (so all threads take the
// same decisions).  Two LDS buffers alternate so consecutive reductions need
// only the one barrier inside each.
template <int K, int M>
__device__ inline void block_reduce(double (&v)[K + M], double* red, int& sel) {
  for (int k = 0; k < K; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  for (int k = K; k < K + M; ++k)
    for (int o = 32; o > 0; o >>= 1) {
      double t = __shfl_xor(v[k], o, 64);
      v[k] = t > v[k] ? t : v[k];
    }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* buf = red + sel * (8 * kMaxWaves);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < K + M; ++k) buf[k * kMaxWaves + w] = v[k];
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += buf[k * kMaxWaves + q];
    v[k] = s;
  }
  for (int k = K; k < K + M; ++k) {
    double s = buf[k * kMaxWaves];
    for (int q = 1; q < nw; ++q) s = buf[k * kMaxWaves + q] > s ? buf[k * kMaxWaves + q] : s;
    v[k] = s;
  }
  sel ^= 1;
}

// Wavefront reductions of FP64 values through DPP lane moves (no LDS).  A
// DPP control moves each 32-bit half: quad_perm [1,0,3,2] (0xB1) and
// [2,3,0,1] (0x4E), row_half_mirror (0x141), row_mirror (0x140).  After the
// four butterfly steps every lane of a 16-lane row holds the row's total,
// formed by the same additions in every lane (IEEE addition commutes), and
// the four row totals are combined in a fixed order: every lane of the wave
// gets the same bits.  All 64 lanes must be active.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}

__device__ inline double readlane_f64(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

__device__ inline double row_sum_f64(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

__device__ inline double row_max_f64(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x141>(v));
  v = fmax(v, dpp_f64<0x140>(v));
  return v;
}

__device__ inline double wave_sum_f64(double v) {
  v = row_sum_f64(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ inline double wave_max_f64(double v) {
  v = row_max_f64(v);
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Block-wide reduction of K sums then M maxima with DPP: each wave reduces
// its own lanes, lane 0 writes the wave's totals to LDS, and after one
// barrier every wave combines the (at most 16) wave totals in lanes 0..15
// with the same row butterfly, so every thread gets the same bits.  Two LDS
// buffers alternate (sel), so consecutive reductions need only the barrier
// inside each.
template <int K, int M>
__device__ inline void block_reduce_dpp(double (&v)[K + M], double* red, int& sel) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < K; ++k) v[k] = wave_sum_f64(v[k]);
  for (int k = K; k < K + M; ++k) v[k] = wave_max_f64(v[k]);
  double* buf = red + sel * (8 * kMaxWaves);
  if (lane == 0)
    for (int k = 0; k < K + M; ++k) buf[k * kMaxWaves + w] = v[k];
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    const double x = lane < nw ? buf[k * kMaxWaves + lane] : 0.0;
    v[k] = readlane_f64(row_sum_f64(x), 0);
  }
  for (int k = K; k < K + M; ++k) {
    const double x = lane < nw ? buf[k * kMaxWaves + lane] : -INFINITY;
    v[k] = readlane_f64(row_max_f64(x), 0);
  }
  sel ^= 1;
}

__device__ inline unsigned nb_index(unsigned e) { return e & 0xFFFFFFu; }
__device__ inline float nb_mult(unsigned e) { return (float)(e >> 24); }

// The workgroup(s) that solve one problem.  SoloTeam: one workgroup.
struct SoloTeam {
  __device__ int rank() const { return 0; }
  __device__ int size() const { return 1; }
  __device__ bool leader() const { return true; }
  __device__ void barrier() { __syncthreads(); }
  template <int K, int M>
  __device__ void reduce(double (&v)[K + M], double* red, int& sel) { block_reduce<K, M>(v, red, sel); }
};

// GridTeam: the problem's vertices (and half-edges) spread over `c`
// co-resident workgroups (item i belongs to workgroup (i / T) % c).  A
// reduction is the workgroup's own block_reduce, written to a per-problem
// slot in global memory, a team barrier, then the c partials summed in rank
// order by every workgroup, so all of them hold the same bits.  The barrier
// is an arrival counter per problem (monotonic; workgroup w waits for
// (n + 1) c arrivals at its n-th barrier) with agent-scope release/acquire
// fences around it: the team's workgroups may sit on different XCDs, whose
// L2 caches are not coherent with each other.  Every wait is bounded by
// `limit` device clock ticks; a wait that runs out sets the shared abort
// flag, after which no barrier waits (the solve finishes in bounded time
// and the host reports the error).
constexpr int kMaxTeam = 64;

struct GridTeam {
  int r, c;
  unsigned* cnt;        // this pair's arrival counter
  double* part;         // this pair's partials [2][kMaxTeam][8]
  int* abort_flag;
  long long limit;
  unsigned arrivals;
  int gsel;
  __device__ int rank() const { return r; }
  __device__ int size() const { return c; }
  __device__ bool leader() const { return r == 0; }
  __device__ void barrier() {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned target = (arrivals + 1u) * (unsigned)c;
      const long long t0 = wall_clock64();
      while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > limit) {
          __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __threadfence();
    }
    ++arrivals;
    __syncthreads();
  }
  template <int K, int M>
  __device__ void reduce(double (&v)[K + M], double* red, int& sel) {
    block_reduce<K, M>(v, red, sel);
    exchange<K, M>(v);
  }
  // The same with the workgroup's DPP reduction.
  template <int K, int M>
  __device__ void reduce_dpp(double (&v)[K + M], double* red, int& sel) {
    block_reduce_dpp<K, M>(v, red, sel);
    exchange<K, M>(v);
  }
  // Team part of a reduction: every workgroup's total (identical in all its
  // threads) goes to its rank's slot; after the barrier each workgroup sums
  // the c slots in rank order.
  template <int K, int M>
  __device__ void exchange(double (&v)[K + M]) {
    double* slot = part + gsel * (kMaxTeam * 8);
    gsel ^= 1;
    if (threadIdx.x == 0)
      for (int k = 0; k < K + M; ++k) slot[r * 8 + k] = v[k];
    barrier();
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int q = 0; q < c; ++q) t += slot[q * 8 + k];
      v[k] = t;
    }
    for (int k = K; k < K + M; ++k) {
      double t = slot[k];
      for (int q = 1; q < c; ++q) t = slot[q * 8 + k] > t ? slot[q * 8 + k] : t;
      v[k] = t;
    }
  }
};

// Device-coherent accesses.  Relaxed agent-scope atomics compile to loads
// and stores with the device-scope cache policy (sc1 on gfx950): they skip
// the CU's L1 and are coherent across the XCDs' L2 caches, so data moved
// through them needs no L2 write-back (buffer_wbl2) or invalidation
// (buffer_inv) -- the cache maintenance an agent-scope fence performs on the
// whole L2 of its XCD, which dominates a fenced team barrier.
__device__ inline void store_coherent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_coherent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_coherent(float4* p, float4 v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  const unsigned long long a = (unsigned long long)__float_as_uint(v.x) | ((unsigned long long)__float_as_uint(v.y) << 32);
  const unsigned long long b = (unsigned long long)__float_as_uint(v.z) | ((unsigned long long)__float_as_uint(v.w) << 32);
  __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float4 load_coherent(const float4* p) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(const_cast<float4*>(p));
  const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float4(__uint_as_float((unsigned)a), __uint_as_float((unsigned)(a >> 32)), __uint_as_float((unsigned)b),
                     __uint_as_float((unsigned)(b >> 32)));
}

// CoherentTeam: GridTeam's protocol (per-problem arrival counter, FP64
// partials in rank slots, bounded waits, shared abort flag; the slots are
// combined by one wave-wide DPP reduction, the same in every member) for
// teams whose exchanged data -- the partials and whatever the
// caller publishes -- goes only through store_coherent / load_coherent.
// Its barrier needs no fence: each wave waits for its own stores to
// complete (s_waitcnt vmcnt(0)), the workgroup barrier then orders them
// before thread 0's arrival, and a coherent load issued after the barrier
// that ends the wait reads the stored value.  Plain global data a member
// reads is only data it wrote itself.
struct CoherentTeam {
  int r, c;
  unsigned* cnt;        // this problem's arrival counter
  double* part;         // this problem's partials [2][kMaxTeam][8]
  int* abort_flag;
  long long limit;
  unsigned arrivals;
  int gsel;
  __device__ bool leader() const { return r == 0; }
  __device__ void barrier() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned target = (arrivals + 1u) * (unsigned)c;
      const long long t0 = wall_clock64();
      while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > limit) {
          __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    ++arrivals;
    __syncthreads();
  }
  template <int K, int M>
  __device__ void reduce_dpp(double (&v)[K + M], double* red, int& sel) {
    block_reduce_dpp<K, M>(v, red, sel);
    double* slot = part + gsel * (kMaxTeam * 8);
    gsel ^= 1;
    if (threadIdx.x == 0)
      for (int k = 0; k < K + M; ++k) store_coherent(slot + r * 8 + k, v[k]);
    barrier();
    // lane q of every wave loads rank q's partials (c <= kMaxTeam = 64), all
    // loads in flight at once; the wave's DPP reduction then gives every
    // lane of every member the same bits
    const int lane = threadIdx.x & 63;
    double x[K + M];
    for (int k = 0; k < K + M; ++k) x[k] = lane < c ? load_coherent(slot + lane * 8 + k) : (k < K ? 0.0 : -INFINITY);
    for (int k = 0; k < K; ++k) v[k] = wave_sum_f64(x[k]);
    for (int k = K; k < K + M; ++k) v[k] = wave_max_f64(x[k]);
  }
};

// Candidate evaluation of one trust-region step.
struct LMCandidate {
  double model_change;   // model_cost_change = -(J delta).(f + J delta / 2)
  double cost;           // cost at x + delta
  double step_norm;      // ||delta||
  double x_norm;         // ||x||
};

// The trust-region loop of Ceres' TrustRegionMinimizer with the
// Levenberg-Marquardt strategy (restated; Ceres is not available here):
//   eval(first, cost, gmax)          residuals, J^T f, cost at x (first: also
//                                    the Jacobi scaling)
//   solve(reuse_diagonal, radius)    LM step delta (returns PCG iterations)
//   candidate()                      -> LMCandidate
//   accept()                         x <- x + delta
// Writes the per-pair result from thread 0 when out is not null (a
// multi-workgroup solve passes it to its first workgroup only).
template <class Eval, class Solve, class Cand, class Accept>
__device__ inline void lm_trust_region_loop(const LMOptionsDev opt, Eval eval, Solve solve, Cand candidate,
                                            Accept accept, LMResultDev* out, long long clock_start) {
  double cv, ce, gmax;
  eval(true, cv, ce, gmax);
  double cost = cv + ce;
  const double initial_cost = cost;
  double radius = opt.initial_trust_region_radius, decrease = 2.0;
  bool reuse = false, last_success = true;
  int iter = 0, invalid = 0, nsucc = 0, nfail = 0, term = 1;
  long long pcg_total = 0;
  while (true) {
    if (iter >= opt.max_num_iterations) { term = 1; break; }
    if (last_success && gmax <= opt.gradient_tolerance) { term = 0; break; }
    if (radius <= opt.min_trust_region_radius) { term = 0; break; }
    ++iter;
    pcg_total += solve(reuse, radius);
    reuse = true;
    const LMCandidate c = candidate();
    if (!(c.model_change > 0.0) || !(c.cost < INFINITY)) {
      if (++invalid >= opt.max_consecutive_invalid_steps) { term = 2; break; }
      radius /= decrease;
      decrease *= 2.0;
      last_success = false;
      ++nfail;
      continue;
    }
    invalid = 0;
    if (c.step_norm <= opt.parameter_tolerance * (c.x_norm + opt.parameter_tolerance)) { term = 0; break; }
    if (fabs(cost - c.cost) <= opt.function_tolerance * cost) { term = 0; break; }
    const double rho = (cost - c.cost) / c.model_change;
    if (rho > opt.min_relative_decrease) {
      accept();
      eval(false, cv, ce, gmax);
      cost = cv + ce;
      const double t = 2.0 * rho - 1.0;
      const double f = 1.0 - t * t * t;
      radius = radius / (f > 1.0 / 3.0 ? f : 1.0 / 3.0);
      radius = radius < opt.max_trust_region_radius ? radius : opt.max_trust_region_radius;
      decrease = 2.0;
      reuse = false;
      last_success = true;
      ++nsucc;
      if (opt.max_successful_steps > 0 && nsucc >= opt.max_successful_steps) { term = 0; break; }
    } else {
      radius /= decrease;
      decrease *= 2.0;
      reuse = true;
      last_success = false;
      ++nfail;
    }
  }
  if (threadIdx.x == 0 && out) {
    LMResultDev r;
    r.initial_cost = initial_cost;
    r.final_cost = cv + ce;
    r.vertex_cost = cv;
    r.rigidity_cost = ce;
    r.iterations = iter;
    r.successful_steps = nsucc;
    r.unsuccessful_steps = nfail;
    r.termination = term;
    r.pcg_iterations = pcg_total;
    r.clock_start = clock_start;
    r.clock_end = wall_clock64();
    *out = r;
  }
}

}  // namespace mo
