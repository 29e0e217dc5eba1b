// LinearEstimation (linear.cc:11-112) and LinearEstimationWithRot
// (linear.cc:114-230): sparse SPD solves on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mo {

// V [nv,3] FP64 device array: rest positions in, solution out.  F [nf,3],
// E [ne,2], ref [nv] and GV [ngv,3] are device arrays.  Returns MO_OK or an
// error; *iterations receives the CG iteration count.
int linear_estimation(double* V, int64_t nv, const int32_t* F, int64_t nf, const int32_t* E, int64_t ne,
                      const int32_t* ref, const double* GV, int64_t ngv, double rigidity, int max_iter, double rtol,
                      int* iterations, hipStream_t s);

// LinearEstimationWithRot (linear.cc:114-230): V [nv,3] FP64 device array,
// rest positions in, solution out; F [nf,3] and TV [nv,3] (target
// positions) device arrays.  Vertices joined by a zero-length face edge
// (where the reference's weight 2e-2/|e| is infinite) share one unknown.
int linear_estimation_rot(double* V, int64_t nv, const int32_t* F, int64_t nf, const double* TV, double rigidity,
                          int max_iter, double rtol, int* iterations, hipStream_t s);

}  // namespace mo
