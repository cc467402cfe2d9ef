#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abf {
__global__ void dist_mat_kernel(const double *gpos, const double *mpos, int64_t G, int M, int64_t g0, int vec2,
                                double *rm);
__global__ void transfer_kernel(const double *r0, const double *rm, int64_t G, int M, double k, int phase_mode,
                                int steer_mode, int out_c128, void *out);
} // namespace abf
