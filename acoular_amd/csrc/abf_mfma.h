// Helpers of the MFMA beamformer kernels (abf_bf.hip).
#pragma once
#include "abf_common.h"
#include "abf_kernels.h"

namespace abf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

// Pointers read from job descriptors in memory are generic (flat) to the compiler; flat
// loads count against the LDS wait counter too, so an LDS wait would also wait for every
// prefetch in flight.  These casts make such accesses global loads/stores.
#define ABF_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ const ABF_GLOBAL T *gp(const T *p) { return (const ABF_GLOBAL T *)p; }
template <typename T> __device__ __forceinline__ ABF_GLOBAL T *gpw(T *p) { return (ABF_GLOBAL T *)p; }

__device__ __forceinline__ int upper_index(int I, int J, int nK) { return I * nK - (I * (I - 1)) / 2 + (J - I); }

// The wavenumbers of a band of at most 64 frequencies f0 .. f1 - 1, lane i holding that of
// frequency f0 + i, read where needed with readlane (a uniform frequency index).  A vector load of
// the wavenumber inside the frequency loop of the streamed kernels would make its first use wait
// for every U load in flight.  (In the alternating kernels, whose loop starts with no loads in
// flight, the same change measured within noise: 59.94-59.97 vs 60.01-60.13 ms on the fused
// 512-mic workload.)
struct BandWavenumbers {
    double v;
    int f0;
    __device__ __forceinline__ BandWavenumbers(const double *wavenumber, int f0_, int f1, int lane) : f0(f0_)
    {
        v = f0_ + lane < f1 ? gp(wavenumber)[f0_ + lane] : 0.0;
        // used here, so the load is waited for once, where the band starts, and not at the top of
        // every frequency (where the compiler cannot tell the first iteration from the others)
        asm volatile("" ::"v"(v));
    }
    __device__ __forceinline__ double operator()(int f) const
    {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, f - f0);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), f - f0);
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    }
};

__device__ __forceinline__ int find_job(const MfJob *__restrict__ jobs, int njobs, int64_t wg)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].wg_begin <= wg)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// x = hi + lo with hi, lo bf16 (round to nearest even on the hardware converter), 8 values
__device__ __forceinline__ void split8(const float *x, u4 &hi, u4 &lo)
{
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const bf16x2 h = {(__bf16)x[2 * d], (__bf16)x[2 * d + 1]};
        const bf16x2 l = {(__bf16)(x[2 * d] - (float)h[0]), (__bf16)(x[2 * d + 1] - (float)h[1])};
        hi[d] = __builtin_bit_cast(uint32_t, h);
        lo[d] = __builtin_bit_cast(uint32_t, l);
    }
}

// x = hi + lo with hi, lo fp16 (round to nearest even), 8 values
__device__ __forceinline__ void split8_h(const float *x, u4 &hi, u4 &lo)
{
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const f16x2 h = {(_Float16)x[2 * d], (_Float16)x[2 * d + 1]};
        const f16x2 l = {(_Float16)(x[2 * d] - (float)h[0]), (_Float16)(x[2 * d + 1] - (float)h[1])};
        hi[d] = __builtin_bit_cast(uint32_t, h);
        lo[d] = __builtin_bit_cast(uint32_t, l);
    }
}

// 8 values rounded to fp16 (round to nearest even)
__device__ __forceinline__ u4 pack8_h(const float *x)
{
    u4 v;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const f16x2 h = {(_Float16)x[2 * d], (_Float16)x[2 * d + 1]};
        v[d] = __builtin_bit_cast(uint32_t, h);
    }
    return v;
}

__device__ __forceinline__ float f16_to_float(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// frequency-independent steering normalisation sum of one microphone (fastFuncs.py:262-273)
__device__ __forceinline__ double help_term(int steer_type, double r)
{
    if (steer_type == ABF_STEER_INVERSE) return r * r;
    if (steer_type == ABF_STEER_TRUE_LEVEL || steer_type == ABF_STEER_TRUE_LOCATION) return 1.0 / (r * r);
    return 0.0;
}

} // namespace abf
