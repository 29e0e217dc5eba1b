// LinearEstimation (linear.cc:11-112): sparse SPD solve on the GPU.
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

}  // namespace mo
