// Time-domain source synthesis for gfx950: the microphone signals of a stationary point source,
// per-channel white noise, and weighted mixing, each computed on the device with the
// reference's arithmetic.
//
// Point source (PointSource.result, acoular/sources.py:774-842, _fill_mic_signal_block
// :77-91): microphone m at output sample b reads the upsampled signal at
// int(0.5 + ind_m(b)) and divides by r_m, where ind_m(0) = (-r_m/c - start_t + start) * fs * up
// and ind_m(b + 1) = ind_m(b) + up, accumulated in float64 one sample at a time.  Repeated float64
// additions round differently from ind_m(0) + b * up once the magnitude crosses a power of two,
// so the accumulation is reproduced exactly: one thread per microphone walks the sequence and
// records its value at the start of every chunk of samples, then one thread per (chunk,
// microphone) continues from there.  Negative indices wrap like numpy's (prepadding 'loop') or
// give 0 ('zeros').
//
// White noise (WNoiseGenerator.signal, acoular/signals.py:255-269, as used per channel by
// UncorrelatedNoiseSource.result, sources.py:1911-1967): numpy's legacy
// RandomState(seed).standard_normal -- the MT19937 generator of Matsumoto & Nishimura (1998)
// seeded with Knuth's multiplier 1812433253, doubles formed from two 32-bit outputs
// ((a >> 5) * 2^26 + (b >> 6)) / 2^53, and the polar Box-Muller method returning f*x2 first and
// f*x1 on the next call.  One 256-thread workgroup per channel keeps the 624-word state in LDS,
// twists it in three parallel phases, tempers a whole batch of outputs, and evaluates all (at
// most 156) candidate pairs of a twist at once: candidates consume the output stream four words
// at a time whether accepted or not, so the waves' ballots and a prefix count place each accepted
// pair's two normals exactly where the sequential generator puts them.  Products and sums are written without fused
// multiply-add (as numpy's baseline x86-64 build computes them; contraction is off for this
// file).  log(r2) is taken as
// log1p(r2 - 1) for r2 >= 1/2 (r2 - 1 is exact there), which keeps the device library's result
// close to glibc's correctly rounded log also where log(r2) is small; the two still differ in
// the last bit for about 1 % of the pairs, whose normals then differ by one ulp.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/abf.h"
#include "abf_status.h"

// every product and sum in this file is rounded on its own, as the reference's host code rounds
// it: hipcc would otherwise fuse a * b + c into one fused multiply-add (also through the
// __dmul_rn / __dadd_rn helpers, whose header bodies allow contraction, so plain operators are
// used under this pragma)
#pragma clang fp contract(off)

namespace abf {

constexpr int kPsChunk = 32;      // samples per thread of the fill kernel
constexpr int64_t kGridY = 65535; // chunks per grid row (grid y), further rows in grid z

// x after n sequential additions x = x + up, each rounded as the reference rounds it.  While the
// partial sums stay inside the binade of x (same sign and exponent) and up is a multiple of that
// binade's ulp (up an integer whose lowest set bit is at least the ulp), every partial sum is a
// representable multiple of the ulp, so the k additions equal the single addition x + k up (k up
// is an exact integer).  Only the addition that leaves the binade can round; it is done as one
// ordinary step.  A run of 51,200 samples thus takes about 35 steps instead of 51,200.
__device__ __forceinline__ double ps_advance(double x, double up, double lowbit, int64_t n)
{
    while (n > 0) {
        double kd = 0.0;
        if (lowbit > 0.0 && x != 0.0 && isfinite(x)) {
            const int e = ilogb(x); // |x| in [2^e, 2^(e+1))
            if (ldexp(1.0, e - 52) <= lowbit) {
                if (x > 0.0) {
                    const double lim = ldexp(1.0, e + 1);
                    kd = floor((lim - x) / up);
                    kd = fmin(kd, (double)n);
                    while (kd > 0.0 && x + kd * up >= lim) kd -= 1.0;
                } else {
                    const double lo = -ldexp(1.0, e);
                    kd = floor((lo - x) / up);
                    kd = fmin(kd, (double)n);
                    while (kd > 0.0 && x + kd * up > lo) kd -= 1.0;
                }
            }
        }
        if (kd >= 1.0) {
            x = x + kd * up;
            n -= (int64_t)kd;
        } else {
            x = x + up;
            n -= 1;
        }
    }
    return x;
}

// start index of every (chunk, microphone): chunk c begins after c * kPsChunk additions
__global__ __launch_bounds__(256) void ps_ind_kernel(const double *__restrict__ ind0, int M, int64_t T, double up,
                                                     int64_t nchunks, double *__restrict__ starts)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y + (int64_t)blockIdx.z * kGridY;
    if (m >= M || c >= nchunks) return;
    // lowest set bit of an integer up (0: not an integer, every addition is taken one by one)
    double lowbit = 0.0;
    if (up >= 1.0 && up < 9007199254740992.0 && up == floor(up)) {
        const uint64_t u = (uint64_t)up;
        lowbit = (double)(u & (~u + 1));
    }
    starts[c * M + m] = ps_advance(ind0[m], up, lowbit, c * kPsChunk);
}

__global__ __launch_bounds__(256) void ps_fill_kernel(const double *__restrict__ signal, int64_t L,
                                                      const double *__restrict__ rm, const double *__restrict__ starts,
                                                      int M, int64_t T, double up, int zeros, double *__restrict__ out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y + (int64_t)blockIdx.z * kGridY;
    if (m >= M || c * kPsChunk >= T) return;
    double x = starts[c * M + m];
    const double r = rm[m];
    const int64_t b0 = c * kPsChunk, b1 = min<int64_t>(T, b0 + kPsChunk);
    for (int64_t b = b0; b < b1; ++b) {
        double v;
        if (zeros && x < 0.0) {
            v = 0.0;
        } else {
            int64_t k = (int64_t)(0.5 + x); // truncation toward zero, as Python's int()
            if (k < 0) k += L;                      // numpy's negative indices
            // an index numpy would reject yields NaN (the caller reports it as an IndexError)
            v = (k >= 0 && k < L) ? signal[k] / r : __builtin_nan("");
        }
        out[b * M + m] = v;
        x = x + up;
    }
}

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMtA = 0x9908b0dfu, kMtUpper = 0x80000000u, kMtLower = 0x7fffffffu;

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ uint32_t mt_next(uint32_t cur, uint32_t nxt, uint32_t far)
{
    const uint32_t y = (cur & kMtUpper) | (nxt & kMtLower);
    return far ^ (y >> 1) ^ ((y & 1u) ? kMtA : 0u);
}

constexpr int kWnT = 256; // threads per stream: one step per twist phase, one candidate round per twist

// one twist of the 624-word state by the kWnT threads of a workgroup: kk in [0, 227) reads old
// words only, [227, 454) reads words the first phase renewed, [454, 624) words the second phase
// renewed (and the last word the renewed word 0)
__device__ void mt_twist(uint32_t *mt, int tid)
{
    uint32_t v = 0;
    if (tid < kMtN - kMtM) v = mt_next(mt[tid], mt[tid + 1], mt[tid + kMtM]);
    __syncthreads();
    if (tid < kMtN - kMtM) mt[tid] = v;
    __syncthreads();
    for (int base = kMtN - kMtM; base < kMtN; base += kMtN - kMtM) {
        const int kk = base + tid;
        const bool on = tid < kMtN - kMtM && kk < kMtN;
        if (on) v = mt_next(mt[kk], mt[(kk + 1) % kMtN], mt[kk - (kMtN - kMtM)]);
        __syncthreads();
        if (on) mt[kk] = v;
        __syncthreads();
    }
}

__device__ __forceinline__ double mt_double(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

// One workgroup per channel (stream).  After each twist the 624 tempered words (after the < 4
// words left over from the previous twist) form candidates of 4 words = two uniform doubles;
// every thread tests one candidate with the polar method and the accepted pairs are placed in
// candidate order through a workgroup prefix of the waves' ballots, so the output is numpy's
// sequence whatever the thread count.
__global__ __launch_bounds__(kWnT) void wnoise_kernel(const uint32_t *__restrict__ seeds, int M, int64_t N, double rms,
                                                      double *__restrict__ out)
{
    __shared__ uint32_t mt[kMtN];
    __shared__ uint32_t words[kMtN + 4];
    __shared__ int s_cnt[kWnT / 64];
    const int ch = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) {
        uint32_t s = seeds[ch];
        for (int pos = 0; pos < kMtN; ++pos) {
            mt[pos] = s;
            s = 1812433253u * (s ^ (s >> 30)) + (uint32_t)(pos + 1);
        }
    }
    __syncthreads();
    int64_t produced = 0;
    int left = 0; // words carried over from the previous twist (< 4)
    while (produced < N) {
        mt_twist(mt, tid);
        for (int i = tid; i < kMtN; i += kWnT) words[left + i] = mt_temper(mt[i]);
        __syncthreads();
        const int total = left + kMtN, ncand = total / 4; // ncand <= 156 < kWnT: one round
        const int c = tid;
        bool acc = false;
        double f = 0.0, x1 = 0.0, x2 = 0.0;
        if (c < ncand) {
            x1 = 2.0 * mt_double(words[4 * c], words[4 * c + 1]) - 1.0;
            x2 = 2.0 * mt_double(words[4 * c + 2], words[4 * c + 3]) - 1.0;
            const double r2 = x1 * x1 + x2 * x2;
            acc = !(r2 >= 1.0 || r2 == 0.0);
            // log(r2) keeps its relative accuracy near r2 = 1 as log1p of the exact r2 - 1
            const double lg = r2 >= 0.5 ? log1p(r2 - 1.0) : log(r2);
            if (acc) f = sqrt(-2.0 * lg / r2);
        }
        const uint64_t mask = __ballot(acc);
        if (lane == 0) s_cnt[wv] = __popcll(mask);
        __syncthreads();
        int64_t before = 0, all = 0;
#pragma unroll
        for (int w = 0; w < kWnT / 64; ++w) {
            before += w < wv ? s_cnt[w] : 0;
            all += s_cnt[w];
        }
        if (acc) {
            const int64_t rank = before + __popcll(mask & ((1ull << lane) - 1ull));
            const int64_t pos = produced + 2 * rank;
            if (pos < N) out[pos * M + ch] = rms * (f * x2);
            if (pos + 1 < N) out[(pos + 1) * M + ch] = rms * (f * x1);
        }
        produced += 2 * all;
        const int rest = total - 4 * ncand;
        const uint32_t keep = tid < rest ? words[4 * ncand + tid] : 0u;
        __syncthreads();
        if (tid < rest) words[tid] = keep;
        left = rest;
        __syncthreads();
    }
}

// acc = w * src (first) or acc + w * src, each product and sum rounded separately
__global__ __launch_bounds__(256) void mix_kernel(const double *__restrict__ src, double w, int64_t n, int first,
                                                  double *__restrict__ acc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p = src[i] * w;
    acc[i] = first ? p : acc[i] + p;
}

} // namespace abf

using namespace abf;

extern "C" {

size_t abf_point_source_workspace_size(int32_t num_mics, int64_t num_samples)
{
    if (num_mics < 1 || num_samples < 0) return 0;
    const int64_t nchunks = (num_samples + kPsChunk - 1) / kPsChunk;
    return (size_t)std::max<int64_t>(nchunks, 1) * num_mics * sizeof(double);
}

int abf_point_source(const double *signal, int64_t signal_len, const double *ind0, const double *rm, int32_t num_mics,
                     int32_t up, int32_t zeros_prepadding, int64_t num_samples, double *out, void *workspace,
                     size_t workspace_bytes, void *stream)
{
    if (num_mics < 1 || num_samples < 0 || signal_len < 1 || up < 1) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_samples == 0) return ABF_OK;
    if (!signal || !ind0 || !rm || !out) return set_error(ABF_ERR_INVALID, "null pointer");
    if (!workspace || workspace_bytes < abf_point_source_workspace_size(num_mics, num_samples))
        return set_error(ABF_ERR_WORKSPACE, "workspace too small");
    const int64_t nchunks = (num_samples + kPsChunk - 1) / kPsChunk;
    if (nchunks > kGridY * 65535) return set_error(ABF_ERR_INVALID, "too many samples");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *starts = static_cast<double *>(workspace);
    const unsigned mb = (unsigned)((num_mics + 255) / 256);
    const dim3 grid(mb, (unsigned)std::min<int64_t>(nchunks, kGridY), (unsigned)((nchunks + kGridY - 1) / kGridY));
    hipLaunchKernelGGL(ps_ind_kernel, grid, dim3(256), 0, s, ind0, num_mics, num_samples, (double)up, nchunks,
                       starts);
    hipLaunchKernelGGL(ps_fill_kernel, grid, dim3(256), 0, s, signal, signal_len, rm, starts, num_mics, num_samples,
                       (double)up, zeros_prepadding != 0, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

int abf_white_noise(const uint32_t *seeds, int32_t num_channels, int64_t num_samples, double rms, double *out,
                    void *stream)
{
    if (num_channels < 0 || num_samples < 0) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_channels == 0 || num_samples == 0) return ABF_OK;
    if (!seeds || !out) return set_error(ABF_ERR_INVALID, "null pointer");
    hipLaunchKernelGGL(wnoise_kernel, dim3((unsigned)num_channels), dim3(kWnT), 0, reinterpret_cast<hipStream_t>(stream),
                       seeds, num_channels, num_samples, rms, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

int abf_mix(const double *src, double weight, int64_t n, int32_t first, double *acc, void *stream)
{
    if (n < 0) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (n == 0) return ABF_OK;
    if (!src || !acc) return set_error(ABF_ERR_INVALID, "null pointer");
    hipLaunchKernelGGL(mix_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), src, weight, n, first, acc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
