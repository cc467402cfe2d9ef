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
    int logP; // log2 P (welch_fft_radix_kernel: P is a power of two)
    int bpw;  // welch_fft_radix_kernel: consecutive blocks per workgroup
    int nbl;  // welch_fft_radix_kernel: blocks of this launch (bl0 ... bl0 + nbl - 1)
    int bl0; // first block of this launch within the pass (blocks b0 + bl0 ...)
};

// workgroups per FFT launch: 2^22 x 512 work-items stays below the 2^32 a dispatch can count
constexpr int64_t kFftMaxWorkgroups = int64_t(1) << 22;

// Workgroups are handed to the eight XCDs in turn (workgroup w runs on XCD w % 8), and each XCD
// has its own L2.  xcd_order(w, n) renumbers the n workgroups of a launch so that XCD x takes one
// contiguous range of logical indices, x q + min(x, r) + w / 8 with n = 8 q + r: the neighbouring
// channel groups of a block (which share the 128-byte lines of every time sample and spectrum
// row) and the next, half-overlapping block then meet in one L2 instead of each XCD fetching
// whole lines for a quarter of their bytes.  A bijection of [0, n).
__device__ __forceinline__ int64_t xcd_order(int64_t w, int64_t n)
{
    const int64_t q = n >> 3, r = n & 7, x = w & 7;
    return x * q + (x < r ? x : r) + (w >> 3);
}

// LDS of welch_fft_radix_kernel for 2^logn-point blocks and p channel pairs: the twiddle table
// (n / 2 entries) and p sequences of n points, indices padded by one slot per 16, the sequence
// stride 8 mod 32 complex slots
__host__ __device__ constexpr int welch_fft_table_len(int n) { return (n / 2 + n / 32 + 31) & ~31; }
__host__ __device__ constexpr int welch_fft_pair_stride(int n) { return ((n + n / 16 + 31) & ~31) + 8; }
inline size_t welch_fft_lds_bytes(int n, int p)
{
    return (size_t)(welch_fft_table_len(n) + (size_t)p * welch_fft_pair_stride(n)) * 8u;
}
// block sizes the radix kernel covers, its threads per workgroup (P bs <= 16 threads) and its
// LDS ceiling (one workgroup per CU)
constexpr int kFftRadixMinLog = 4, kFftRadixMaxLog = 12;
constexpr int kFftRadixThreads = 1024;
constexpr size_t kFftRadixLdsMax = 160 * 1024;

__global__ void fill_zero_kernel(float4 *p, int64_t n4);
__global__ void calc_csm_c64_kernel(float2 *csm, const float2 *spec, int f0, int M, int vec2);
__global__ void calc_csm_c128_kernel(double2 *csm, const double2 *spec, int f0, int M);
__global__ void welch_fft_kernel(WelchFftArgs a);
template <int LOGN, bool VEC> __global__ void welch_fft_radix_kernel(WelchFftArgs a);
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
