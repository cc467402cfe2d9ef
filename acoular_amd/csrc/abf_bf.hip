// Frequency-domain beamformer kernels for gfx950 (MI355X).
//
// B(g, f) = h_g^H C_f h_g for every grid point g and frequency slice f, with the steering
// vector h generated on the fly from grid and microphone coordinates
// (acoular/fastFuncs.py:238-311).  Every kernel reads the CSM through the "upper-weighted"
// matrix U of one frequency:
//     U_ij = 2 C_ij (j > i),   U_ii = Re C_ii (0 when the diagonal is removed),   U_ij = 0 (j < i)
// so that  q = Re sum_i conj(v_i) (U v)_i  equals the reference's Hermitian half-sum
// 2 Re sum_{i<j} C_ij conj(v_i) v_j + sum_i C_ii |v_i|^2 (fastFuncs.py:277-291), reading only
// the upper triangle and the diagonal of C like the reference does.
//
// Two forms:
//  * plain FP32 (bf_plain_kernel): one grid point per lane, U streamed as wave-uniform
//    (scalar) loads, the steering vectors of 64 grid points staged in LDS per mic chunk;
//  * MFMA complex GEMM (bf_mfma_kernel): Y = U X on 16x16x32 bf16 MFMA with both operands
//    split hi+lo (3 products per real product) or on 16x16x4 fp32 MFMA; X (the steering
//    matrix of a grid tile) lives in LDS, U is pre-laid-out in MFMA fragment order by
//    prep_mfma_kernel and streamed from L2; q = Re sum_i conj(x_i) Y_i in the epilogue.
#include "abf_common.h"
#include "abf_kernels.h"

namespace abf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

// -------------------------------------------------------------------------------------------
// steering element: v = w * (cos phi, -sin phi), phi = k r
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ void phasor_rev(double phi, int phase_mode, int fast, float &c, float &s)
{
    if (phase_mode == ABF_PHASE_REFERENCE) phi = (double)(float)phi;
    if (fast) {
        // hardware sin/cos take their argument in revolutions; reduce to [-1/2, 1/2] in float64
        const double u = phi * kInvTwoPi;
        const float t = (float)(u - rint(u));
        s = __builtin_amdgcn_sinf(t);
        c = __builtin_amdgcn_cosf(t);
    } else {
        const double n = rint(phi * kInvTwoPi);
        double t = fma(-n, kTwoPiHi, phi);
        t = fma(-n, kTwoPiLo, t);
        sincosf((float)t, &s, &c);
    }
}

// -------------------------------------------------------------------------------------------
// U^T for the plain kernel: ut[f][j][i] = U_ij, zero padded to Mp x Mp
// -------------------------------------------------------------------------------------------
__global__ void prep_plain_kernel(const float2 *__restrict__ csm, int M, int Mp, int r_diag, float2 *__restrict__ ut)
{
    const int f = blockIdx.y;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)Mp * Mp) return;
    const int j = (int)(idx / Mp), i = (int)(idx % Mp);
    float2 v = make_float2(0.f, 0.f);
    if (i < M && j < M) {
        const float2 c = csm[((int64_t)f * M + i) * M + j];
        if (j > i)
            v = make_float2(2.f * c.x, 2.f * c.y);
        else if (j == i && !r_diag)
            v = make_float2(c.x, 0.f);
    }
    ut[((int64_t)f * Mp + j) * Mp + i] = v;
}

// -------------------------------------------------------------------------------------------
// plain FP32 kernel
// -------------------------------------------------------------------------------------------
constexpr int kPlainNP = 64;   // grid points per workgroup (one per lane)
constexpr int kPlainRB = 16;   // rows of U per register block
constexpr int kPlainJC = 256;  // microphones per LDS chunk
constexpr int kPlainWaves = 4;

struct PlainPoint {
    double gx, gy, gz;
    int64_t g;
    bool valid;
};

__device__ __forceinline__ float2 plain_steer(const PlainArgs &a, const PlainPoint &p, int m, double k, double &help)
{
    help = 0.0;
    if (!p.valid || m >= a.M) return make_float2(0.f, 0.f);
    double r;
    if (a.rm_table)
        r = a.rm_table[p.g * a.M + m];
    else
        r = dist3(p.gx, p.gy, p.gz, a.mpos[m], a.mpos[a.M + m], a.mpos[2 * a.M + m]);
    float c, s, w;
    phasor_rev(k * r, a.phase_mode, a.fast_trig, c, s);
    steer_weight(a.steer_type, r, w, help);
    return make_float2(w * c, -w * s);
}

__global__ __launch_bounds__(256) void bf_plain_kernel(PlainArgs a)
{
    extern __shared__ float2 vlds[]; // [kPlainJC][kPlainNP]
    __shared__ double s_help[kPlainWaves][kPlainNP];
    __shared__ float s_q[kPlainWaves][kPlainNP];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int f = blockIdx.y;
    PlainPoint p;
    p.g = (int64_t)blockIdx.x * kPlainNP + lane;
    p.valid = p.g < a.G;
    p.gx = p.gy = p.gz = 0.0;
    if (p.valid && !a.rm_table) {
        p.gx = a.gpos[p.g];
        p.gy = a.gpos[a.G + p.g];
        p.gz = a.gpos[2 * a.G + p.g];
    }
    const double k = a.wavenumber[f];
    const float2 *__restrict__ ut = a.ut + (int64_t)f * a.Mp * a.Mp;
    double help = 0.0;
    float q = 0.f;

    for (int c0 = 0; c0 < a.Mp; c0 += kPlainJC) {
        const int jn = min(kPlainJC, a.Mp - c0);
        __syncthreads();
        for (int mm = wave; mm < jn; mm += kPlainWaves) {
            double h;
            vlds[mm * kPlainNP + lane] = plain_steer(a, p, c0 + mm, k, h);
            help += h;
        }
        __syncthreads();
        const int nrb = (c0 + jn) / kPlainRB;
        const int rounds = (nrb + kPlainWaves - 1) / kPlainWaves;
        for (int rd = 0; rd < rounds; ++rd) {
            // boustrophedon assignment keeps the triangular work balanced over the waves
            const int rb = rd * kPlainWaves + ((rd & 1) ? (kPlainWaves - 1 - wave) : wave);
            if (rb >= nrb) continue;
            const int i0 = rb * kPlainRB;
            const int js = max(i0, c0);
            float ar[kPlainRB], ai[kPlainRB];
#pragma unroll
            for (int r = 0; r < kPlainRB; ++r) ar[r] = ai[r] = 0.f;
            for (int j = js; j < c0 + jn; ++j) {
                const float2 vj = vlds[(j - c0) * kPlainNP + lane];
                const float2 *__restrict__ u = ut + (int64_t)j * a.Mp + i0;
#pragma unroll
                for (int r = 0; r < kPlainRB; ++r) {
                    const float2 uu = u[r];
                    ar[r] = fmaf(uu.x, vj.x, ar[r]);
                    ar[r] = fmaf(-uu.y, vj.y, ar[r]);
                    ai[r] = fmaf(uu.x, vj.y, ai[r]);
                    ai[r] = fmaf(uu.y, vj.x, ai[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < kPlainRB; ++r) {
                const int i = i0 + r;
                float2 vi;
                if (i >= c0) {
                    vi = vlds[(i - c0) * kPlainNP + lane];
                } else {
                    double h;
                    vi = plain_steer(a, p, i, k, h);
                }
                q = fmaf(vi.x, ar[r], q);
                q = fmaf(vi.y, ai[r], q);
            }
        }
    }
    s_q[wave][lane] = q;
    s_help[wave][lane] = help;
    __syncthreads();
    if (wave == 0 && p.valid) {
        double qs = 0.0, hs = 0.0;
#pragma unroll
        for (int w = 0; w < kPlainWaves; ++w) {
            qs += (double)s_q[w][lane];
            hs += s_help[w][lane];
        }
        double r0;
        if (a.r0_table)
            r0 = a.r0_table[p.g];
        else if (a.ref_dist > 0.0)
            r0 = a.ref_dist;
        else
            r0 = dist3(p.gx, p.gy, p.gz, a.ref[0], a.ref[1], a.ref[2]);
        double res, nst;
        normalise(a.steer_type, qs, r0, hs, (double)a.M, a.norm, res, nst);
        float rf = (float)res;
        if (a.clip && rf < 0.f) rf = 0.f;
        a.out[(int64_t)f * a.G + p.g] = rf;
        if (a.out_norm) a.out_norm[(int64_t)f * a.G + p.g] = (float)nst;
    }
}

// -------------------------------------------------------------------------------------------
// MFMA form
// -------------------------------------------------------------------------------------------
// Fragment layout of U for one frequency: upper 32x32 blocks (I <= J) in row-major upper
// order; per block two 16-row tiles; per tile 4 KiB:
//   bf16x3: parts {Ur_hi, Ur_lo, Ui_hi, Ui_lo}, each [64 lanes][8 bf16]
//   f32:    parts {Ur, Ui}, each [64 lanes][8 f32]
// Lane l of a tile holds row (l & 15) and columns 8*(l >> 4) + 0..7 of the block, the same
// k-set the B operand (X) uses, so each MFMA pairs matching k.
__device__ __forceinline__ int upper_index(int I, int J, int nK) { return I * nK - (I * (I - 1)) / 2 + (J - I); }

__global__ __launch_bounds__(64) void prep_mfma_kernel(const float2 *__restrict__ csm, int M, int Mp, int r_diag,
                                                       int bf16x3, char *__restrict__ prep)
{
    const int nK = Mp / 32;
    const int I = blockIdx.x / nK, J = blockIdx.x % nK;
    if (J < I) return;
    const int t = blockIdx.y, f = blockIdx.z;
    const int l = threadIdx.x;
    const int i = I * 32 + t * 16 + (l & 15);
    const int jb = J * 32 + (l >> 4) * 8;
    float ur[8], ui[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j = jb + e;
        float x = 0.f, y = 0.f;
        if (i < M && j < M) {
            const float2 c = csm[((int64_t)f * M + i) * M + j];
            if (j > i) {
                x = 2.f * c.x;
                y = 2.f * c.y;
            } else if (j == i && !r_diag) {
                x = c.x;
            }
        }
        ur[e] = x;
        ui[e] = y;
    }
    const int nUp = nK * (nK + 1) / 2;
    char *dst = prep + (((int64_t)f * nUp + upper_index(I, J, nK)) * 2 + t) * 4096;
    if (bf16x3) {
        uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            uint32_t h0, l0, h1, l1;
            split_bf16(ur[e], h0, l0);
            split_bf16(ur[e + 1], h1, l1);
            rh[e / 2] = h0 | (h1 << 16);
            rl[e / 2] = l0 | (l1 << 16);
            split_bf16(ui[e], h0, l0);
            split_bf16(ui[e + 1], h1, l1);
            ih[e / 2] = h0 | (h1 << 16);
            il[e / 2] = l0 | (l1 << 16);
        }
        u4 *d = reinterpret_cast<u4 *>(dst);
        d[0 * 64 + l] = u4{rh[0], rh[1], rh[2], rh[3]};
        d[1 * 64 + l] = u4{rl[0], rl[1], rl[2], rl[3]};
        d[2 * 64 + l] = u4{ih[0], ih[1], ih[2], ih[3]};
        d[3 * 64 + l] = u4{il[0], il[1], il[2], il[3]};
    } else {
        f4 *d = reinterpret_cast<f4 *>(dst);
        d[(0 * 64 + l) * 2 + 0] = f4{ur[0], ur[1], ur[2], ur[3]};
        d[(0 * 64 + l) * 2 + 1] = f4{ur[4], ur[5], ur[6], ur[7]};
        d[(1 * 64 + l) * 2 + 0] = f4{ui[0], ui[1], ui[2], ui[3]};
        d[(1 * 64 + l) * 2 + 1] = f4{ui[4], ui[5], ui[6], ui[7]};
    }
}

constexpr int kMfWaves = 8;
constexpr int kMfThreads = kMfWaves * 64;
constexpr int kMfNS = 5; // max 8-mic groups per thread in the steering prologue

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

template <bool BF16X3, int CW>
__global__ __launch_bounds__(kMfThreads, 1) void bf_mfma_kernel(const MfJob *__restrict__ jobs, int njobs)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t wg = blockIdx.x;
    const int jidx = find_job(jobs, njobs, wg);
    const MfJob &jb = jobs[jidx];

    const int M = jb.M, Mp = jb.Mp, CT = jb.CT, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int BN = 16 * CT;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int fgrp = (int)(local / jb.ntiles);
    const int f0 = fgrp * FB;
    const int f1 = min(f0 + FB, F);
    const int KS = BF16X3 ? Mp + 8 : Mp + 4; // LDS row stride in elements (16-byte skew)
    const int nK = Mp / 32;
    const int nUp = nK * (nK + 1) / 2;

    // LDS carve-up
    uint16_t *xh16 = reinterpret_cast<uint16_t *>(smem);     // bf16x3: 4 planes [BN][KS]
    float *x32 = reinterpret_cast<float *>(smem);            // f32: 2 planes [BN][KS]
    const int64_t plane_bytes = BF16X3 ? (int64_t)BN * KS * 2 : (int64_t)BN * KS * 4;
    const int nplanes = BF16X3 ? 4 : 2;
    double *s_r0 = reinterpret_cast<double *>(smem + plane_bytes * nplanes);
    double *s_help = s_r0 + BN;
    float *s_q = reinterpret_cast<float *>(s_help + BN); // [kMfWaves][BN]

    // ---- frequency-independent prologue: distances kept in registers ----
    const int TPC = kMfThreads / BN; // threads per grid column (power of two)
    const int col = tid / TPC, part = tid % TPC;
    const int64_t g = (int64_t)tile * BN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = jb.gpos[g];
        gy = jb.gpos[G + g];
        gz = jb.gpos[2 * G + g];
    }
    const int ngroups = Mp / 8;
    double rr[kMfNS][8];
    double help = 0.0;
#pragma unroll
    for (int s = 0; s < kMfNS; ++s) {
        const int gi = part + TPC * s;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = gi * 8 + e;
            double r = 1.0;
            if (gi < ngroups && valid && m < M) {
                r = dist3(gx, gy, gz, jb.mpos[m], jb.mpos[M + m], jb.mpos[2 * M + m]);
                float w;
                double h;
                steer_weight(jb.steer_type, r, w, h);
                help += h;
            }
            rr[s][e] = r;
        }
    }
    for (int off = TPC >> 1; off > 0; off >>= 1) help += __shfl_xor(help, off);
    if (part == 0) {
        s_help[col] = help;
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
        s_r0[col] = r0;
    }
    for (int idx = tid; idx < kMfWaves * BN; idx += kMfThreads) s_q[idx] = 0.f;

    const int RT = Mp / 16;
    const int NCG = CT / CW;
    const int nunits = RT * NCG;
    const int rounds = (nunits + kMfWaves - 1) / kMfWaves;
    const char *__restrict__ prep_base = reinterpret_cast<const char *>(jb.prep);

    for (int f = f0; f < f1; ++f) {
        const double k = jb.wavenumber[f];
        // ---- steering matrix X of this frequency into LDS ----
#pragma unroll
        for (int s = 0; s < kMfNS; ++s) {
            const int gi = part + TPC * s;
            if (gi < ngroups) {
                float xr[8], xi[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int m = gi * 8 + e;
                    float c, sn, w;
                    double h;
                    phasor_rev(k * rr[s][e], jb.phase_mode, jb.fast_trig, c, sn);
                    steer_weight(jb.steer_type, rr[s][e], w, h);
                    const bool live = valid && m < M;
                    xr[e] = live ? w * c : 0.f;
                    xi[e] = live ? -w * sn : 0.f;
                }
                const int64_t off = (int64_t)col * KS + gi * 8;
                if constexpr (BF16X3) {
                    uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
                    for (int e = 0; e < 8; e += 2) {
                        uint32_t h0, l0, h1, l1;
                        split_bf16(xr[e], h0, l0);
                        split_bf16(xr[e + 1], h1, l1);
                        rh[e / 2] = h0 | (h1 << 16);
                        rl[e / 2] = l0 | (l1 << 16);
                        split_bf16(xi[e], h0, l0);
                        split_bf16(xi[e + 1], h1, l1);
                        ih[e / 2] = h0 | (h1 << 16);
                        il[e / 2] = l0 | (l1 << 16);
                    }
                    const int64_t pe = (int64_t)BN * KS; // plane stride in elements
                    *reinterpret_cast<u4 *>(xh16 + off) = u4{rh[0], rh[1], rh[2], rh[3]};
                    *reinterpret_cast<u4 *>(xh16 + pe + off) = u4{rl[0], rl[1], rl[2], rl[3]};
                    *reinterpret_cast<u4 *>(xh16 + 2 * pe + off) = u4{ih[0], ih[1], ih[2], ih[3]};
                    *reinterpret_cast<u4 *>(xh16 + 3 * pe + off) = u4{il[0], il[1], il[2], il[3]};
                } else {
                    const int64_t pe = (int64_t)BN * KS;
                    *reinterpret_cast<f4 *>(x32 + off) = f4{xr[0], xr[1], xr[2], xr[3]};
                    *reinterpret_cast<f4 *>(x32 + off + 4) = f4{xr[4], xr[5], xr[6], xr[7]};
                    *reinterpret_cast<f4 *>(x32 + pe + off) = f4{xi[0], xi[1], xi[2], xi[3]};
                    *reinterpret_cast<f4 *>(x32 + pe + off + 4) = f4{xi[4], xi[5], xi[6], xi[7]};
                }
            }
        }
        __syncthreads();

        // ---- Y = U X over the upper block triangle, q = Re sum conj(x) Y ----
        const char *__restrict__ prep_f = prep_base + (int64_t)f * nUp * 8192;
        for (int rd = 0; rd < rounds; ++rd) {
            const int u = rd * kMfWaves + ((rd & 1) ? (kMfWaves - 1 - wave) : wave);
            if (u >= nunits) continue;
            const int rt = u / NCG, cg = u % NCG;
            const int I = rt >> 1, t = rt & 1;
            f4 accr[CW], acci[CW];
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                accr[c] = f4{0.f, 0.f, 0.f, 0.f};
                acci[c] = f4{0.f, 0.f, 0.f, 0.f};
            }
            const char *ablk = prep_f + ((int64_t)upper_index(I, I, nK) * 2 + t) * 4096;
            const int colbase = cg * CW * 16 + (lane & 15);
            const int kq = (lane >> 4) * 8;
            for (int J = I; J < nK; ++J, ablk += 8192) {
                const int koff = J * 32 + kq;
                if constexpr (BF16X3) {
                    const u4 *ap = reinterpret_cast<const u4 *>(ablk);
                    const u4 a_rh = ap[0 * 64 + lane], a_rl = ap[1 * 64 + lane];
                    const u4 a_ih = ap[2 * 64 + lane], a_il = ap[3 * 64 + lane];
                    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
                    const bf16x8 Arh = __builtin_bit_cast(bf16x8, a_rh), Arl = __builtin_bit_cast(bf16x8, a_rl);
                    const bf16x8 Aih = __builtin_bit_cast(bf16x8, a_ih), Ail = __builtin_bit_cast(bf16x8, a_il);
                    const bf16x8 Nih = __builtin_bit_cast(bf16x8, a_ih ^ sg), Nil = __builtin_bit_cast(bf16x8, a_il ^ sg);
                    const int64_t pe = (int64_t)BN * KS;
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        const int64_t xo = (int64_t)(colbase + c * 16) * KS + koff;
                        const bf16x8 Xrh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + xo));
                        const bf16x8 Xrl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + pe + xo));
                        const bf16x8 Xih = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + 2 * pe + xo));
                        const bf16x8 Xil = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + 3 * pe + xo));
                        // Yr = Ur Xr - Ui Xi ; Yi = Ui Xr + Ur Xi  (each real product as hi*hi + hi*lo + lo*hi)
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arh, Xrh, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Aih, Xrh, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arh, Xrl, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Aih, Xrl, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arl, Xrh, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ail, Xrh, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Nih, Xih, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arh, Xih, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Nih, Xil, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arh, Xil, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Nil, Xih, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Arl, Xih, acci[c], 0, 0, 0);
                    }
                } else {
                    const f4 *ap = reinterpret_cast<const f4 *>(ablk);
                    const f4 ur0 = ap[(0 * 64 + lane) * 2], ur1 = ap[(0 * 64 + lane) * 2 + 1];
                    const f4 ui0 = ap[(1 * 64 + lane) * 2], ui1 = ap[(1 * 64 + lane) * 2 + 1];
                    const float ur[8] = {ur0[0], ur0[1], ur0[2], ur0[3], ur1[0], ur1[1], ur1[2], ur1[3]};
                    const float ui[8] = {ui0[0], ui0[1], ui0[2], ui0[3], ui1[0], ui1[1], ui1[2], ui1[3]};
                    const int64_t pe = (int64_t)BN * KS;
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        const int64_t xo = (int64_t)(colbase + c * 16) * KS + koff;
                        const f4 xr0 = *reinterpret_cast<const f4 *>(x32 + xo);
                        const f4 xr1 = *reinterpret_cast<const f4 *>(x32 + xo + 4);
                        const f4 xi0 = *reinterpret_cast<const f4 *>(x32 + pe + xo);
                        const f4 xi1 = *reinterpret_cast<const f4 *>(x32 + pe + xo + 4);
                        const float xr[8] = {xr0[0], xr0[1], xr0[2], xr0[3], xr1[0], xr1[1], xr1[2], xr1[3]};
                        const float xi[8] = {xi0[0], xi0[1], xi0[2], xi0[3], xi1[0], xi1[1], xi1[2], xi1[3]};
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[e], xr[e], accr[c], 0, 0, 0);
                            acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[e], xr[e], acci[c], 0, 0, 0);
                            accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui[e], xi[e], accr[c], 0, 0, 0);
                            acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[e], xi[e], acci[c], 0, 0, 0);
                        }
                    }
                }
            }
            // epilogue: lane holds rows rt*16 + 4*(lane>>4) + 0..3 of column colbase + c*16
            const int row0 = rt * 16 + (lane >> 4) * 4;
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int colc = colbase + c * 16;
                const int64_t xo = (int64_t)colc * KS + row0;
                float sacc = 0.f;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float xr, xi;
                    if constexpr (BF16X3) {
                        const int64_t pe = (int64_t)BN * KS;
                        xr = bf16_to_float(xh16[xo + e]) + bf16_to_float(xh16[pe + xo + e]);
                        xi = bf16_to_float(xh16[2 * pe + xo + e]) + bf16_to_float(xh16[3 * pe + xo + e]);
                    } else {
                        const int64_t pe = (int64_t)BN * KS;
                        xr = x32[xo + e];
                        xi = x32[pe + xo + e];
                    }
                    sacc = fmaf(xr, accr[c][e], sacc);
                    sacc = fmaf(xi, acci[c][e], sacc);
                }
                sacc += __shfl_xor(sacc, 16);
                sacc += __shfl_xor(sacc, 32);
                if (lane < 16) s_q[wave * BN + colc] += sacc;
            }
        }
        __syncthreads();

        // ---- normalise and store ----
        for (int cidx = tid; cidx < BN; cidx += kMfThreads) {
            double qs = 0.0;
#pragma unroll
            for (int w = 0; w < kMfWaves; ++w) {
                qs += (double)s_q[w * BN + cidx];
                s_q[w * BN + cidx] = 0.f;
            }
            const int64_t gg = (int64_t)tile * BN + cidx;
            if (gg < G) {
                double res, nst;
                normalise(jb.steer_type, qs, s_r0[cidx], s_help[cidx], (double)M, jb.norm, res, nst);
                float rf = (float)res;
                if (jb.clip && rf < 0.f) rf = 0.f;
                jb.out[(int64_t)f * G + gg] = rf;
                if (jb.out_norm) jb.out_norm[(int64_t)f * G + gg] = (float)nst;
            }
        }
        // the next frequency overwrites X only after every wave has passed this barrier
        __syncthreads();
    }
}

template __global__ void bf_mfma_kernel<true, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<true, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<true, 4>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 4>(const MfJob *, int);

// -------------------------------------------------------------------------------------------
// eigenvalue form (fastFuncs.py:335-411) and explicit-steering forms (fastFuncs.py:414-559):
// one grid point per thread, eigenvectors / CSM read as wave-uniform loads.
// -------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bf_eig_kernel(EigArgs a)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = g < a.G;
    const int64_t gg = valid ? g : 0;
    double q = 0.0, help = 0.0;
    // help sum over microphones (table or explicit steering)
    for (int e = 0; e < a.K; ++e) {
        float tr = 0.f, ti = 0.f, dsum = 0.f;
        for (int m = 0; m < a.M; ++m) {
            float2 v;
            if (a.steer) {
                v = a.steer[gg * a.M + m];
            } else {
                const double r = a.rm_table[gg * a.M + m];
                float c, s, w;
                double h;
                phasor_rev(a.wavenumber * r, a.phase_mode, a.fast_trig, c, s);
                steer_weight(a.steer_type, r, w, h);
                v = make_float2(w * c, -w * s);
            }
            const float2 ev = a.eigvec_t[(int64_t)e * a.M + m]; // conj(V_me) * v_m
            const float pr = ev.x * v.x + ev.y * v.y;
            const float pi = ev.x * v.y - ev.y * v.x;
            tr += pr;
            ti += pi;
            dsum = fmaf(pr, pr, fmaf(pi, pi, dsum));
        }
        double full = (double)tr * tr + (double)ti * ti;
        if (a.r_diag) full -= dsum;
        q += full * a.eigval[e];
    }
    for (int m = 0; m < a.M; ++m) {
        if (a.steer) {
            const float2 v = a.steer[gg * a.M + m];
            help += (double)v.x * v.x + (double)v.y * v.y;
        } else {
            const double r = a.rm_table[gg * a.M + m];
            float w;
            double h;
            steer_weight(a.steer_type, r, w, h);
            help += h;
        }
    }
    if (!valid) return;
    double res, nst;
    if (a.steer) {
        res = q * a.norm;
        nst = help;
    } else {
        normalise(a.steer_type, q, a.r0_table[g], help, (double)a.M, a.norm, res, nst);
    }
    a.out[g] = (float)res;
    if (a.out_norm) a.out_norm[g] = (float)nst;
}

__global__ __launch_bounds__(256) void bf_custom_kernel(CustomArgs a)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.G) return;
    const float2 *h = a.steer + g * a.M;
    float q = 0.f, help = 0.f;
    for (int j = 0; j < a.M; ++j) {
        const float2 hj = h[j];
        help = fmaf(hj.x, hj.x, fmaf(hj.y, hj.y, help));
        float lr = 0.f, li = 0.f; // sum_{i<j} C_ij conj(h_i)
        for (int i = 0; i < j; ++i) {
            const float2 c = a.csm[(int64_t)i * a.M + j];
            const float2 hi = h[i];
            lr = fmaf(c.x, hi.x, fmaf(c.y, hi.y, lr));
            li = fmaf(c.y, hi.x, fmaf(-c.x, hi.y, li));
        }
        q = fmaf(2.f, lr * hj.x - li * hj.y, q);
        if (!a.r_diag) {
            const float cd = a.csm[(int64_t)j * a.M + j].x;
            q = fmaf(cd, hj.x * hj.x + hj.y * hj.y, q);
        }
    }
    a.out[g] = (float)((double)q * a.norm);
    if (a.out_norm) a.out_norm[g] = help;
}

} // namespace abf
