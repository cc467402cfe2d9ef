#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abf {
// grid points per workgroup of dist_mat_kernel (131,072 x 1024 distances: 0.346 ms with 1, 0.239
// with 4, 0.226 with 8, 0.207-0.209 with 16 = 5.2 TB/s, 0.217 with 32, 0.234 with 64)
#ifndef ABF_DIST_PTS
#define ABF_DIST_PTS 16
#endif
constexpr int kDistPts = ABF_DIST_PTS;
__global__ void dist_mat_kernel(const double *gpos, const double *mpos, int64_t G, int M, int64_t g0, int vec2,
                                double *rm);
__global__ void transfer_kernel(const double *r0, const double *rm, int64_t G, int M, double k, int phase_mode,
                                int steer_mode, int out_c128, void *out);
} // namespace abf
