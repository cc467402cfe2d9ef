// Windowed block FFT of PowerSpectra.calc_csm's Welch estimate (acoular/spectra.py:366-416: the
// window times each block, then the one-sided FFT per channel) and of fprocess.RFFT, for block
// sizes 16 ... 4096: Stockham autosort passes of radix 8 or 16 whose butterflies run in registers.
//
// Two real channels (a, b) of one block travel as one complex sequence z = a + i b; after the
// transform X_a[k] = (Z[k] + conj(Z[N-k])) / 2 and X_b[k] = (Z[k] - conj(Z[N-k])) / (2 i).
// A workgroup of 1024 threads transforms P channel pairs (P a power of two, P N <= 16384) of a
// run of consecutive blocks:
//   pass 0      each item (sample offset j, pair p) loads the R samples j + r N / R of its two
//               channels straight from the time data (pairs fastest: the 2 P channels of one time
//               sample are adjacent floats, 64 bytes for P = 8), windows them, takes their R-point
//               DFT and writes the outputs to LDS at j R + k;
//   pass q > 0  each item reads R values at stride N / R from LDS into registers, multiplies by
//               the twiddles exp(-2 pi i (j mod Ns) r / (Ns R)), takes the R-point DFT and writes
//               the outputs at (j / Ns) Ns R + (j mod Ns) + k Ns (Ns = product of the earlier
//               radices), so that after the last pass the spectrum is in natural order;
//   output      pairs (k, N - k) are combined into the two channels' spectra and written as one
//               16-byte store per channel pair (channels 2p, 2p + 1 are adjacent in memory).
// LDS holds the P sequences (stride S = 8 mod 32 two-word banks, so the P pairs of one k fall on
// distinct banks) and the twiddle table exp(-2 pi i k / N), k < N / 2; every index is padded by one
// slot per 16 (i + i / 16) so the stride-R writes of the first passes spread over the banks.
// For N = 2048, P = 8 takes 148 KB of LDS: one workgroup of 16 waves per CU, whose 64-byte loads
// and 128-byte stores run at 3.2 TB/s on 512 channels (512 threads and P = 4: 2.86 TB/s; 256
// threads and P = 2: 2.08 TB/s).
#include <type_traits>

#include "abf_common.h"
#include "abf_csm_kernels.h"

namespace abf {

namespace {

constexpr int kFftThreads = kFftRadixThreads;
constexpr int kFftPN = 16 * kFftThreads; // the largest P N: at most 16 points per thread and pass

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }

// bits of pass q for a transform of 2^LOGN points: ceil(LOGN / 4) passes of at most 4 bits, the
// first passes take the larger share
template <int LOGN> struct FftPasses {
    static constexpr int count = (LOGN + 3) / 4;
    static constexpr int bits(int q) { return LOGN / count + (q < LOGN % count ? 1 : 0); }
    static constexpr int before(int q) { return q == 0 ? 0 : before(q - 1) + bits(q - 1); }
};

template <int B> constexpr int rev_bits(int r) { return B == 0 ? 0 : ((r & 1) << (B - 1)) | rev_bits<B - 1>(r >> 1); }
template <> constexpr int rev_bits<0>(int) { return 0; }

// x * exp(-2 pi i m / 16), m = 1 ... 7 (m = 4 is -i)
__device__ __forceinline__ float2 rot16(float2 x, int m)
{
    constexpr float c[8] = {1.0f, 0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f,
                            0.0f, -0.382683432365089772f, -0.707106781186547524f, -0.923879532511286756f};
    constexpr float s[8] = {0.0f, 0.382683432365089772f, 0.707106781186547524f, 0.923879532511286756f,
                            1.0f, 0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f};
    if (m == 4) return make_float2(x.y, -x.x);
    return make_float2(x.x * c[m] + x.y * s[m], x.y * c[m] - x.x * s[m]);
}

// in-register DFT of 2^B points, natural order in and out (radix-2 decimation in time over the
// bit-reversed copy; the twiddles of the 2h-point stage are exp(-2 pi i k / 2h) = W16^(8 k / h))
template <int B> __device__ __forceinline__ void dft_reg(float2 (&v)[1 << B])
{
    constexpr int R = 1 << B;
    float2 t[R];
#pragma unroll
    for (int r = 0; r < R; ++r) t[rev_bits<B>(r)] = v[r];
#pragma unroll
    for (int h = 1; h < R; h <<= 1)
#pragma unroll
        for (int i = 0; i < R; i += 2 * h)
#pragma unroll
            for (int k = 0; k < h; ++k) {
                float2 x = t[i + k + h];
                if (k != 0) x = rot16(x, k * (8 / h));
                const float2 u = t[i + k];
                t[i + k] = make_float2(u.x + x.x, u.y + x.y);
                t[i + k + h] = make_float2(u.x - x.x, u.y - x.y);
            }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = t[r];
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(-2 pi i e / N) for 0 <= e < N from the table of its first half
template <int LOGN> __device__ __forceinline__ float2 twiddle_at(const float2 *tw, int e)
{
    constexpr int H = 1 << (LOGN - 1);
    const float2 w = tw[fpad(e & (H - 1))];
    return (e & H) ? make_float2(-w.x, -w.y) : w;
}

// Barrier for the LDS hand-over between passes: orders LDS accesses only, so that global loads of
// the next block stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// v[r] *= exp(-2 pi i e0 r / N) for 0 < r < 2^B, from the B table entries of e0 2^b: each power
// is the product of the power with its lowest set bit cleared and one table entry, taken in
// increasing r so that few powers are live at once
template <int LOGN, int B> __device__ __forceinline__ void twiddle_apply(const float2 *tw, int e0, float2 (&v)[1 << B])
{
    float2 wp[B];
#pragma unroll
    for (int b = 0; b < B; ++b) wp[b] = twiddle_at<LOGN>(tw, (e0 << b) & ((1 << LOGN) - 1));
    float2 w[1 << B];
#pragma unroll
    for (int r = 1; r < (1 << B); ++r) {
        const int low = r & -r;
        const int b = low == 1 ? 0 : low == 2 ? 1 : low == 4 ? 2 : 3;
        w[r] = r == low ? wp[b] : cmul(w[r - low], wp[b]);
        v[r] = cmul(v[r], w[r]);
    }
}

// one LDS pass q >= 1 (see the file comment).  Every item of a thread has the same j (N / R
// divides the 512 threads).
template <int LOGN, int Q> __device__ __forceinline__ void lds_pass(float2 *z, const float2 *tw, int P, int S)
{
    constexpr int B = FftPasses<LOGN>::bits(Q);
    constexpr int R = 1 << B;
    constexpr int LNS = FftPasses<LOGN>::before(Q); // log2 Ns
    constexpr int NS = 1 << LNS;
    constexpr int NR = 1 << (LOGN - B); // items per sequence
    constexpr int IT = (kFftPN >> B) / kFftThreads > 0 ? (kFftPN >> B) / kFftThreads : 1;
    static_assert(kFftThreads % NR == 0, "items of one thread share j");
    const int items = P * NR;
    const int j = threadIdx.x & (NR - 1);
    // LDS offsets: padding adds i / 16, so a multiple of 16 added to an index moves it by 17 / 16
    const int rd0 = fpad(j);
    const int jm = j & (NS - 1);
    const int d = ((j >> LNS) << (LNS + B)) + jm;
    const int wr0 = fpad(d);
    float2 v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * kFftThreads;
        if (item < items) {
            const float2 *base = z + (item >> (LOGN - B)) * S;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (NR % 16 == 0) v[it][r] = base[rd0 + r * (NR + NR / 16)];
                else v[it][r] = base[fpad(j + r * NR)];
            }
        }
    }
    lds_barrier();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * kFftThreads;
        if (item < items) {
            float2 *base = z + (item >> (LOGN - B)) * S;
            twiddle_apply<LOGN, B>(tw, jm << (LOGN - LNS - B), v[it]);
            dft_reg<B>(v[it]);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if constexpr (NS % 16 == 0) base[wr0 + k * (NS + NS / 16)] = v[it][k];
                else base[fpad(d + k * NS)] = v[it][k];
            }
        }
    }
    lds_barrier();
}

template <int LOGN, int Q> __device__ __forceinline__ void lds_passes(float2 *z, const float2 *tw, int P, int S)
{
    if constexpr (Q < FftPasses<LOGN>::count) {
        lds_pass<LOGN, Q>(z, tw, P, S);
        lds_passes<LOGN, Q + 1>(z, tw, P, S);
    }
}

} // namespace

// grid: (channel groups x chunks of a.bpw consecutive blocks of this launch) workgroups, channel
// group fastest, in the XCD order of xcd_order; kFftThreads threads; dynamic LDS
// welch_fft_lds_bytes(2^LOGN, P).  The window values of a thread's samples stay in registers for
// all its blocks, and the twiddle table is loaded once per workgroup.  VEC: M even, the time
// data 8-byte and the spectra 16-byte aligned, so that a channel pair is one 8-byte load and one
// 16-byte store (ma is even).
template <int LOGN, bool VEC>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(4))) void welch_fft_radix_kernel(WelchFftArgs a)
{
    constexpr int N = 1 << LOGN;
    constexpr int B0 = FftPasses<LOGN>::bits(0);
    constexpr int R0 = 1 << B0;
    constexpr int NR0 = N / R0;
    constexpr int IT0 = (kFftPN / R0) / kFftThreads > 0 ? (kFftPN / R0) / kFftThreads : 1;
    constexpr int F = N / 2 + 1;
    extern __shared__ float2 lds[];
    const int P = a.P, logP = a.logP, M = a.M;
    const int S = welch_fft_pair_stride(N);
    float2 *tw = lds;
    float2 *z = lds + welch_fft_table_len(N);
    const int ng = (M + 2 * P - 1) / (2 * P);
    const int64_t L = xcd_order(blockIdx.x, (int64_t)gridDim.x);
    const int m0 = (int)(L % ng) * 2 * P;
    const int blo = (int)(L / ng) * a.bpw, bhi = min(a.nbl, blo + a.bpw); // blocks bl0 + [blo, bhi)
    for (int k = threadIdx.x; k < N / 2; k += kFftThreads) tw[fpad(k)] = a.twiddle[k];
    // pass-0 items of this thread: (sample offset j, pair p), pairs fastest
    const int items0 = P * NR0;
    const uint32_t rstride = (uint32_t)NR0 * (uint32_t)M; // elements between samples j + r NR0
    bool live[IT0];
    uint32_t off[IT0]; // element offset of sample j, channel m0 + 2 p within a block
    float wv[IT0][R0]; // window / 2 (the 1/2 of the split into two channels), 0 where not live
#pragma unroll
    for (int it = 0; it < IT0; ++it) {
        const int item = threadIdx.x + it * kFftThreads;
        const int p = item & (P - 1), j = item >> logP;
        live[it] = item < items0 && m0 + 2 * p < M;
        off[it] = (uint32_t)j * (uint32_t)M + 2u * p;
#pragma unroll
        for (int r = 0; r < R0; ++r) wv[it][r] = live[it] ? 0.5f * a.window[j + r * NR0] : 0.f;
    }
    const bool two = m0 + 2 * P <= M || !(M & 1); // the second channel of every live pair exists
    float2 xr[IT0][R0];
    // unconditional loads (a lane without a live pair reads the group's first sample, which exists)
    // zeroed afterwards where not live: no branch around the loads, so all of them are in flight
    // together
#pragma unroll
    for (int it = 0; it < IT0; ++it)
        if (!live[it]) off[it] = 0;
    // samples r >= RF of block blk
    auto load = [&](int blk, auto rf) __attribute__((always_inline)) {
        constexpr int RF = decltype(rf)::value;
        const float *tb = a.time + (int64_t)(a.b0 + a.bl0 + blk) * a.hop * M + m0;
        // the offsets are formed here for every block: hoisted out of the block loop they would
        // hold 2 R0 registers per item for the whole kernel
        uint32_t ob[IT0];
#pragma unroll
        for (int it = 0; it < IT0; ++it) {
            ob[it] = off[it];
            asm volatile("" : "+v"(ob[it]));
        }
#pragma unroll
        for (int it = 0; it < IT0; ++it)
#pragma unroll
            for (int r = RF; r < R0; ++r) {
                const uint32_t o = live[it] ? ob[it] + (uint32_t)r * rstride : 0u;
                float2 x;
                if constexpr (VEC) x = *reinterpret_cast<const float2 *>(tb + o);
                else {
                    // the second channel of the last pair exists unless M is odd and this is it
                    const int ma = m0 + 2 * (int)((threadIdx.x + it * kFftThreads) & (P - 1));
                    x.x = tb[o];
                    x.y = tb[(two || ma + 1 < M) ? o + 1 : o];
                    if (!(two || ma + 1 < M)) x.y = 0.f;
                }
                xr[it][r] = x;
            }
    };
    // 50 % overlap (hop = N / 2 = (R0 / 2) NR0): the samples r >= R0 / 2 of a block are the samples
    // r - R0 / 2 of the next one, so a workgroup's blocks after its first load only their second
    // halves and keep the rest in registers (the L2 does not hold them from one block to the next)
    const bool half = 2 * a.hop == N;
    float2 keep[IT0][R0 / 2];
    for (int blk = blo; blk < bhi; ++blk) {
        const int bl = a.bl0 + blk;
        if (half && blk > blo) {
#pragma unroll
            for (int it = 0; it < IT0; ++it)
#pragma unroll
                for (int r = 0; r < R0 / 2; ++r) xr[it][r] = keep[it][r];
            load(blk, std::integral_constant<int, R0 / 2>{});
        } else {
            load(blk, std::integral_constant<int, 0>{});
        }
        float2 v[IT0][R0];
#pragma unroll
        for (int it = 0; it < IT0; ++it)
#pragma unroll
            for (int r = 0; r < R0; ++r) // a lane without a live pair has w = 0
                v[it][r] = make_float2(xr[it][r].x * wv[it][r], xr[it][r].y * wv[it][r]);
#pragma unroll
        for (int it = 0; it < IT0; ++it)
#pragma unroll
            for (int r = 0; r < R0 / 2; ++r) keep[it][r] = xr[it][r + R0 / 2];
#pragma unroll
        for (int it = 0; it < IT0; ++it) {
            const int item = threadIdx.x + it * kFftThreads;
            if (item < items0) {
                dft_reg<B0>(v[it]);
                float2 *base = z + (item & (P - 1)) * S + fpad((item >> logP) * R0); // R0 divides 16
#pragma unroll
                for (int k = 0; k < R0; ++k) base[k] = v[it][k];
            }
        }
        lds_barrier();
        lds_passes<LOGN, 1>(z, tw, P, S);
        // one-sided spectra of both channels of each pair
        float2 *dst0 = a.block_major ? a.spec + (int64_t)(a.b0 + bl) * F * M : a.spec + (int64_t)bl * M;
        const int64_t kstride = a.block_major ? (int64_t)M : (int64_t)a.nb * M;
        for (int idx = threadIdx.x; idx < P * F; idx += kFftThreads) {
            const int p = idx & (P - 1), k = idx >> logP;
            const int ma = m0 + 2 * p;
            if (ma >= M) continue;
            const float2 *base = z + p * S;
            const float2 zk = base[fpad(k)];
            const float2 zn = base[fpad((N - k) & (N - 1))];
            // (the window carries the factor 1/2 of the split into the two channels)
            float2 va = make_float2(zk.x + zn.x, zk.y - zn.y);
            float2 vb = make_float2(zk.y + zn.y, zn.x - zk.x);
            if (a.bin_scale) {
                const float sc = a.bin_scale[k];
                va = make_float2(va.x * sc, va.y * sc);
                vb = make_float2(vb.x * sc, vb.y * sc);
            }
            float2 *dst = dst0 + k * kstride + ma;
            if constexpr (VEC) *reinterpret_cast<float4 *>(dst) = make_float4(va.x, va.y, vb.x, vb.y);
            else {
                dst[0] = va;
                if (ma + 1 < M) dst[1] = vb;
            }
        }
        lds_barrier(); // the next block's pass 0 overwrites the sequences
    }
}

template __global__ void welch_fft_radix_kernel<4, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<4, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<5, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<5, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<6, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<6, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<7, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<7, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<8, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<8, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<9, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<9, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<10, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<10, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<11, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<11, true>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<12, false>(WelchFftArgs);
template __global__ void welch_fft_radix_kernel<12, true>(WelchFftArgs);

} // namespace abf
