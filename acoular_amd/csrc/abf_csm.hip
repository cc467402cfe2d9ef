// Cross-spectral matrix kernels for gfx950.
//
//  * calc_csm_kernel: the in-place upper-triangle rank-1 update of calcCSM
//    (acoular/fastFuncs.py:30-58), complex64 or complex128.
//  * Welch CSM (PowerSpectra.calc_csm, acoular/spectra.py:366-416):
//      welch_fft_kernel      windowed real FFT of every (block, channel), two channels packed
//                            into one complex radix-2 FFT in LDS;
//      welch_herk_kernel     C_f += X_f X_f^H over the blocks of a pass on fp32 MFMA
//                            (16x16x4), upper 64x64 tiles only;
//      welch_finalize_kernel Hermitian completion from the upper triangle with a real
//                            diagonal (spectra.py:412-414) and the 2/(bs w num_blocks) scaling
//                            (spectra.py:416).
#include "abf_common.h"
#include "abf_csm_kernels.h"

namespace abf {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void fill_zero_kernel(float4 *__restrict__ p, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// calcCSM update, csm[f][r][c] += conj(spec[f][c]) spec[f][r] for r <= c.  grid: (ceil(M / 512)
// column blocks, M rows, frequencies f0 ... of this launch): two adjacent columns per thread, as
// one 16-byte access when M is even (every row and spectrum then starts on an even element);
// indices come from the grid, with no 64-bit divisions per element; column pairs left of the
// diagonal exit at once.
__global__ __launch_bounds__(256) void calc_csm_c64_kernel(float2 *__restrict__ csm, const float2 *__restrict__ spec,
                                                           int f0, int M, int vec2)
{
    const int c = (blockIdx.x * 256 + threadIdx.x) * 2, r = blockIdx.y;
    const int64_t f = (int64_t)f0 + blockIdx.z;
    if (c >= M || c + 1 < r) return;
    const float2 sr = spec[f * M + r];
    float2 *row = csm + (f * M + r) * M;
    const float2 *sp = spec + f * M;
    if (vec2 && c >= r) { // both columns on or above the diagonal, c + 1 < M
        float4 v = *reinterpret_cast<const float4 *>(row + c);
        const float4 sc = *reinterpret_cast<const float4 *>(sp + c);
        v.x += sc.x * sr.x + sc.y * sr.y;
        v.y += sc.x * sr.y - sc.y * sr.x;
        v.z += sc.z * sr.x + sc.w * sr.y;
        v.w += sc.z * sr.y - sc.w * sr.x;
        *reinterpret_cast<float4 *>(row + c) = v;
        return;
    }
    for (int cc = c; cc < c + 2 && cc < M; ++cc) {
        if (cc < r) continue;
        const float2 sc = sp[cc];
        float2 v = row[cc];
        v.x += sc.x * sr.x + sc.y * sr.y;
        v.y += sc.x * sr.y - sc.y * sr.x;
        row[cc] = v;
    }
}

__global__ __launch_bounds__(256) void calc_csm_c128_kernel(double2 *__restrict__ csm,
                                                            const double2 *__restrict__ spec, int f0, int M)
{
    const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
    const int64_t f = (int64_t)f0 + blockIdx.z;
    if (c >= M || r > c) return;
    const double2 sr = spec[f * M + r], sc = spec[f * M + c];
    double2 *p = csm + (f * M + r) * M + c;
    double2 v = *p;
    v.x += sc.x * sr.x + sc.y * sr.y;
    v.y += sc.x * sr.y - sc.y * sr.x;
    *p = v;
}

__device__ __forceinline__ int bitrev(int n, int bits) { return (int)(__builtin_bitreverse32((uint32_t)n) >> (32 - bits)); }

// radix-2 decimation-in-time stages over P sequences of bs points in LDS (bit-reversed order on
// entry, natural order on exit), up to three stages per pass through LDS: a thread takes the 2^g
// elements base + j half (j < 2^g) that stages s .. s+g-1 combine only among themselves and runs
// those stages in registers.  Ends with a barrier.
__device__ __forceinline__ void fft_stages(float2 *z, int P, int bs, int logbs, const float2 *twiddle)
{
    for (int s = 1; s <= logbs;) {
        const int g = min(3, logbs - s + 1);
        const int half = 1 << (s - 1);
        const int per_chan = bs >> g; // groups per channel
        for (int idx = threadIdx.x; idx < P * per_chan; idx += blockDim.x) {
            const int p = idx / per_chan, gi = idx % per_chan;
            const int blk = gi / half, pos = gi % half;
            const int base = p * bs + blk * (half << g) + pos;
            float2 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < (1 << g)) v[j] = z[base + j * half];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t >= g) break;
                const int tstride = bs >> (s + t);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (j >= (1 << g) || (j & (1 << t))) continue;
                    const int jj = j | (1 << t);
                    const float2 w = twiddle[(pos + (j & ((1 << t) - 1)) * half) * tstride];
                    const float2 x1 = v[jj];
                    const float tr = x1.x * w.x - x1.y * w.y;
                    const float ti = x1.x * w.y + x1.y * w.x;
                    v[jj] = make_float2(v[j].x - tr, v[j].y - ti);
                    v[j] = make_float2(v[j].x + tr, v[j].y + ti);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < (1 << g)) z[base + j * half] = v[j];
        }
        __syncthreads();
        s += g;
    }
}

// grid: (ceil(M / 2P), nb); P channel pairs of one Welch block per workgroup
__global__ __launch_bounds__(512) void welch_fft_kernel(WelchFftArgs a)
{
    extern __shared__ float2 z[]; // [P][bs]
    const int bs = a.bs, P = a.P;
    // grid: (channel groups x blocks of this launch) workgroups, channel group fastest
    const int ng = (a.M + 2 * P - 1) / (2 * P);
    const int64_t L = xcd_order(blockIdx.x, (int64_t)gridDim.x);
    const int bl = a.bl0 + (int)(L / ng);
    const int m0 = (int)(L % ng) * 2 * P;
    const int64_t t0 = (int64_t)(a.b0 + bl) * a.hop;
    for (int idx = threadIdx.x; idx < P * bs; idx += blockDim.x) {
        const int p = idx % P, n = idx / P;
        const int ma = m0 + 2 * p, mb = ma + 1;
        const float w = a.window[n];
        const float *row = a.time + (t0 + n) * a.M;
        const float xa = ma < a.M ? row[ma] : 0.f;
        const float xb = mb < a.M ? row[mb] : 0.f;
        z[p * bs + bitrev(n, a.logbs)] = make_float2(w * xa, w * xb);
    }
    __syncthreads();
    fft_stages(z, P, bs, a.logbs, a.twiddle);
    const int F = bs / 2 + 1;
    for (int idx = threadIdx.x; idx < P * F; idx += blockDim.x) {
        const int p = idx % P, k = idx / P;
        const float2 zk = z[p * bs + k];
        const float2 zn = z[p * bs + ((bs - k) & (bs - 1))];
        const int ma = m0 + 2 * p, mb = ma + 1;
        float2 *dst = a.block_major ? a.spec + ((int64_t)(a.b0 + bl) * F + k) * a.M
                                    : a.spec + ((int64_t)k * a.nb + bl) * a.M;
        float2 va = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
        float2 vb = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
        if (a.bin_scale) {
            const float sc = a.bin_scale[k];
            va = make_float2(va.x * sc, va.y * sc);
            vb = make_float2(vb.x * sc, vb.y * sc);
        }
        if (ma < a.M) dst[ma] = va;
        if (mb < a.M) dst[mb] = vb;
    }
}

// Writes the upper triangle only (r <= c): the completion reads nothing else.
// grid: nTr * nTc tiles x the frequencies f0 ... of this launch, tile fastest, in the XCD order of
// xcd_order (the tiles of one frequency then read its spectra through one L2); one wave per tile
// of 16 TR rows x 64 columns of one frequency on or above the diagonal (TR = 4: 64 x 64 tiles, 128 accumulator registers; TR = 2: 32 x 64 tiles, 64
// accumulator registers and twice the waves per SIMD, one third more loads per MFMA).  The
// waves-per-EU bound keeps the accumulators in VGPRs so several waves share a SIMD and hide each
// other's load latency; without it the compiler takes 136 VGPRs + 128 AGPRs for TR = 4, one wave
// per SIMD.
template <int TR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8 / TR))) void welch_herk_kernel(
    const float2 *__restrict__ spec, int nb, int M, int nTc, int f0, int accumulate, float2 *__restrict__ csm,
    int nsplit, int F, float2 *__restrict__ part)
{
    const int tiles = (M + 16 * TR - 1) / (16 * TR) * nTc;
    const int64_t L = xcd_order(blockIdx.x, (int64_t)gridDim.x);
    const int tile = (int)(L % tiles);
    const int I = tile / nTc, J = tile % nTc;
    const int r0 = I * 16 * TR, c0 = J * 64;
    if (c0 + 63 < r0) return; // entirely below the diagonal
    const int64_t rest = L / tiles;
    const int sp = (int)(rest % nsplit); // slice of the pass's blocks
    const int f = f0 + (int)(rest / nsplit);
    // blocks [kb0, kb1) of this slice, slices 4-block aligned
    const int chunk = ((nb + nsplit - 1) / nsplit + 3) & ~3;
    const int kb0 = min(nb, sp * chunk), kb1 = min(nb, kb0 + chunk);
    const int l = threadIdx.x;
    const float2 *X = spec + (int64_t)f * nb * M;
    f4 accr[TR][4], acci[TR][4];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            accr[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            acci[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
    // spectra of blocks k0 + (l >> 4): rows r0 + 16 t + (l & 15) and columns c0 + 16 t + (l & 15)
    auto step = [&](const float2 (&ra)[TR], const float2 (&ca)[4]) __attribute__((always_inline)) {
#pragma unroll
        for (int ti = 0; ti < TR; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                // C_rc = sum_b X_r conj(X_c)
                accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[ti].x, ca[tj].x, accr[ti][tj], 0, 0, 0);
                accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[ti].y, ca[tj].y, accr[ti][tj], 0, 0, 0);
                acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[ti].y, ca[tj].x, acci[ti][tj], 0, 0, 0);
                acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ra[ti].x, ca[tj].y, acci[ti][tj], 0, 0, 0);
            }
    };
    if constexpr (TR == 4) {
        // 64 x 64 tiles (three waves per SIMD): the next step's loads are issued before this step's
        // 64 MFMAs and stay in flight across them.  They are unconditional, from an index clamped
        // into the pass (nb >= 1), and zeroed outside it only after the MFMAs: with branches
        // around them, or the zeroing beside them, the compiler waits for them at once.  (For 32-row
        // tiles the extra registers cost a wave per SIMD: 13.9 against 12.4 ms.)
        auto load = [&](int k0, float2 (&ra)[TR], float2 (&ca)[4]) __attribute__((always_inline)) {
            const float2 *row = X + (int64_t)min(k0 + (l >> 4), nb - 1) * M;
#pragma unroll
            for (int t = 0; t < TR; ++t) ra[t] = row[min(r0 + t * 16 + (l & 15), M - 1)];
#pragma unroll
            for (int t = 0; t < 4; ++t) ca[t] = row[min(c0 + t * 16 + (l & 15), M - 1)];
        };
        auto mask = [&](int k0, float2 (&ra)[TR], float2 (&ca)[4]) __attribute__((always_inline)) {
            const bool bv = k0 + (l >> 4) < kb1;
#pragma unroll
            for (int t = 0; t < TR; ++t)
                if (!(bv && r0 + t * 16 + (l & 15) < M)) ra[t] = make_float2(0.f, 0.f);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (!(bv && c0 + t * 16 + (l & 15) < M)) ca[t] = make_float2(0.f, 0.f);
        };
        float2 ra[TR], ca[4];
        load(kb0, ra, ca);
        mask(kb0, ra, ca);
        for (int k0 = kb0; k0 < kb1; k0 += 4) {
            float2 na[TR], nc[4];
            load(k0 + 4, na, nc);
            __builtin_amdgcn_sched_barrier(0);
            step(ra, ca);
            __builtin_amdgcn_sched_barrier(0);
            mask(k0 + 4, na, nc);
#pragma unroll
            for (int t = 0; t < TR; ++t) ra[t] = na[t];
#pragma unroll
            for (int t = 0; t < 4; ++t) ca[t] = nc[t];
        }
    } else {
        // 32 x 64 tiles (five waves per SIMD hide each other's loads)
        for (int k0 = kb0; k0 < kb1; k0 += 4) {
            const int b = k0 + (l >> 4);
            const bool bv = b < kb1;
            float2 ra[TR], ca[4];
#pragma unroll
            for (int t = 0; t < TR; ++t) {
                const int r = r0 + t * 16 + (l & 15);
                ra[t] = (bv && r < M) ? X[(int64_t)b * M + r] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int c = c0 + t * 16 + (l & 15);
                ca[t] = (bv && c < M) ? X[(int64_t)b * M + c] : make_float2(0.f, 0.f);
            }
            step(ra, ca);
        }
    }
    // a split pass stores its slices' partial sums (welch_reduce_kernel adds them)
    float2 *C = nsplit > 1 ? part + ((int64_t)sp * F + f) * M * M : csm + (int64_t)f * M * M;
    const bool add = nsplit == 1 && accumulate;
#pragma unroll
    for (int ti = 0; ti < TR; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = r0 + ti * 16 + (l >> 4) * 4 + i;
                const int c = c0 + tj * 16 + (l & 15);
                if (r < M && c < M && r <= c) {
                    // the first pass stores, later passes add
                    float2 v = add ? C[(int64_t)r * M + c] : make_float2(0.f, 0.f);
                    v.x += accr[ti][tj][i];
                    v.y += acci[ti][tj][i];
                    C[(int64_t)r * M + c] = v;
                }
            }
}

template __global__ void welch_herk_kernel<4>(const float2 *, int, int, int, int, int, float2 *, int, int, float2 *);
template __global__ void welch_herk_kernel<2>(const float2 *, int, int, int, int, int, float2 *, int, int, float2 *);

// Sum of the nsplit slices' partial upper triangles of a split pass, slice by slice in order (the
// same result on every run), stored into the CSM (first pass) or added to it.  grid: (ceil(M /
// 256) column blocks, M rows, F).
__global__ __launch_bounds__(256) void welch_reduce_kernel(const float2 *__restrict__ part, int nsplit, int F, int M,
                                                           int accumulate, float2 *__restrict__ csm)
{
    const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y, f = blockIdx.z;
    if (c >= M || r > c) return;
    const int64_t e = ((int64_t)f * M + r) * M + c, stride = (int64_t)F * M * M;
    float2 v = accumulate ? csm[e] : make_float2(0.f, 0.f);
    for (int s = 0; s < nsplit; ++s) {
        const float2 p = part[s * stride + e];
        v.x += p.x;
        v.y += p.y;
    }
    csm[e] = v;
}

// Hermitian completion (spectra.py:412-414) and scaling (spectra.py:416) by 32 x 32 tiles: tile
// (I, J) on or above the diagonal is read once, scaled and staged in LDS, then written back and
// written conjugate-transposed into tile (J, I), so that every global access is a row segment;
// the diagonal comes out real and the lower half of a diagonal tile is taken from its mirror.
// grid: (nT * nT tiles, F), 256 threads; tiles below the diagonal return at once.
__global__ __launch_bounds__(256) void welch_finalize_kernel(float2 *__restrict__ csm, int M, float scale)
{
    __shared__ float2 t[32][33];
    const int nT = (M + 31) / 32;
    const int I = blockIdx.x / nT, J = blockIdx.x % nT;
    if (I > J) return;
    float2 *C = csm + (int64_t)blockIdx.y * M * M;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int r0 = I * 32, c0 = J * 32;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        float2 v = make_float2(0.f, 0.f);
        if (r < M && c < M) {
            v = C[(int64_t)r * M + c];
            v.x *= scale;
            v.y *= scale;
        }
        t[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < M && c < M) {
            float2 v = t[k][tx];
            if (r == c) {
                v.y = 0.f;
            } else if (r > c) { // only inside a diagonal tile: the conjugate of (c, r)
                const float2 w = t[tx][k];
                v = make_float2(w.x, -w.y);
            }
            C[(int64_t)r * M + c] = v;
        }
    }
    if (I < J) {
        for (int k = ty; k < 32; k += 8) {
            const int r = c0 + k, c = r0 + tx; // element (r, c) of tile (J, I) mirrors (c, r)
            if (r < M && c < M) {
                const float2 w = t[tx][k];
                C[(int64_t)r * M + c] = make_float2(w.x, -w.y);
            }
        }
    }
}

// IRFFT.result (fprocess.py:227-283): the real inverse transform of every snapshot spectrum
// [nb][F][M], two channels per complex sequence: Z = Xa + i Xb over all bs bins (the upper half
// from Hermitian symmetry, the imaginary parts of the DC and Nyquist bins dropped as a real
// inverse transform drops them), z = conj(FFT(conj Z)) / bs = xa + i xb.  Output time-major
// [nb * bs][M].  grid: ceil(M / 2P) channel groups x blocks of this launch, channel group fastest,
// in the XCD order of xcd_order
__global__ __launch_bounds__(512) void irfft_kernel(const float2 *__restrict__ spec, int M, int bs, int logbs, int b0,
                                                    int P, const float2 *__restrict__ twiddle, float *__restrict__ out)
{
    extern __shared__ float2 z[]; // [P][bs]
    const int F = bs / 2 + 1;
    const int ng = (M + 2 * P - 1) / (2 * P);
    const int64_t L = xcd_order(blockIdx.x, (int64_t)gridDim.x);
    const int64_t bl = (int64_t)b0 + L / ng; // b0: first block of this launch
    const int m0 = (int)(L % ng) * 2 * P;
    for (int idx = threadIdx.x; idx < P * bs; idx += blockDim.x) {
        const int p = idx % P, k = idx / P;
        const int ma = m0 + 2 * p, mb = ma + 1;
        const int kk = k < F ? k : bs - k;
        const float2 *row = spec + (bl * F + kk) * M;
        float2 xa = ma < M ? row[ma] : make_float2(0.f, 0.f);
        float2 xb = mb < M ? row[mb] : make_float2(0.f, 0.f);
        if (kk == 0 || kk == F - 1) {
            xa.y = 0.f;
            xb.y = 0.f;
        }
        if (k >= F) { // bin bs - k mirrored: conjugate
            xa.y = -xa.y;
            xb.y = -xb.y;
        }
        // conj(Z) with Z = xa + i xb
        z[p * bs + bitrev(k, logbs)] = make_float2(xa.x - xb.y, -(xa.y + xb.x));
    }
    __syncthreads();
    fft_stages(z, P, bs, logbs, twiddle);
    const float inv = 1.f / (float)bs; // a power of two: exact
    for (int idx = threadIdx.x; idx < P * bs; idx += blockDim.x) {
        const int p = idx % P, n = idx / P;
        const int ma = m0 + 2 * p, mb = ma + 1;
        const float2 v = z[p * bs + n];
        float *dst = out + (bl * bs + n) * M;
        if (ma < M) dst[ma] = v.x * inv;
        if (mb < M) dst[mb] = -v.y * inv;
    }
}

// AutoPowerSpectra.result (fprocess.py:359-366): |X|^2 * scale per value
__global__ __launch_bounds__(256) void auto_power_kernel(const float2 *__restrict__ spec, int64_t n, float scale,
                                                         float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 x = spec[i];
    out[i] = (x.x * x.x + x.y * x.y) * scale;
}

// CrossPowerSpectra.result, calc_mode 'full' (fprocess.py:446-470): per snapshot and bin the
// matrix X X^H, element (r, c) = scale * X_r conj(X_c) (upper part from calcCSM, lower part its
// conjugate transpose); grid (ceil(M^2 / 256), snapshots x bins)
__global__ __launch_bounds__(256) void cross_power_kernel(const float2 *__restrict__ spec, int64_t nbf, int M,
                                                          float scale, float2 *__restrict__ out)
{
    const int64_t sb = blockIdx.y + (int64_t)blockIdx.z * 65535; // snapshot x bin
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= nbf || e >= (int64_t)M * M) return;
    const int r = (int)(e / M), c = (int)(e % M);
    const float2 xr = spec[sb * M + r], xc = spec[sb * M + c];
    out[sb * M * M + e] = make_float2((xr.x * xc.x + xr.y * xc.y) * scale, (xr.y * xc.x - xr.x * xc.y) * scale);
}

} // namespace abf
