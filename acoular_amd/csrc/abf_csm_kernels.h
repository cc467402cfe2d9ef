#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abf {

struct WelchFftArgs {
    const float *time;     // [N][M]
    const float *window;   // [bs]
    const float2 *twiddle; // [bs/2]  exp(-2 pi i j / bs)
    float2 *spec;          // [bs/2+1][nb][M], or [num_blocks][bs/2+1][M] when block_major
    const float *bin_scale; // [bs/2+1] factor per frequency bin, or null
    int64_t N;
    int M, bs, logbs, hop, b0, nb, P, block_major;
    int bl0; // first block of this launch within the pass (blocks b0 + bl0 ...)
};

// workgroups per FFT launch: 2^22 x 512 work-items stays below the 2^32 a dispatch can count
constexpr int64_t kFftMaxWorkgroups = int64_t(1) << 22;

__global__ void fill_zero_kernel(float4 *p, int64_t n4);
__global__ void calc_csm_c64_kernel(float2 *csm, const float2 *spec, int f0, int M, int vec2);
__global__ void calc_csm_c128_kernel(double2 *csm, const double2 *spec, int f0, int M);
__global__ void welch_fft_kernel(WelchFftArgs a);
template <int TR>
__global__ void welch_herk_kernel(const float2 *spec, int nb, int M, int nTc, int f0, int accumulate, float2 *csm,
                                  int nsplit, int F, float2 *part);
__global__ void welch_reduce_kernel(const float2 *part, int nsplit, int F, int M, int accumulate, float2 *csm);
__global__ void welch_finalize_kernel(float2 *csm, int M, float scale);
__global__ void irfft_kernel(const float2 *spec, int M, int bs, int logbs, int b0, int P, const float2 *twiddle,
                             float *out);
__global__ void auto_power_kernel(const float2 *spec, int64_t n, float scale, float *out);
__global__ void cross_power_kernel(const float2 *spec, int64_t nbf, int M, float scale, float2 *out);

} // namespace abf
