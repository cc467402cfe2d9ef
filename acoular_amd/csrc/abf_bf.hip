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
#include <type_traits>

#include "abf_mfma.h"

namespace abf {

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

__global__ __launch_bounds__(64) void prep_mfma_kernel(const float2 *__restrict__ csm, int M, int Mp, int r_diag,
                                                       int bf16x3, char *__restrict__ prep)
{
    const int nK = Mp / 32;
    const int I = blockIdx.x / nK, J = blockIdx.x % nK;
    if (J < I) return;
    const int t = blockIdx.y, f = blockIdx.z;
    const int l = threadIdx.x;
    const int i = I * 32 + t * 16 + (l & 15);
    float ur[8], ui[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j = J * 32 + kslot_mic(l >> 4, e);
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
        u4 rh, rl, ih, il;
        split8(ur, rh, rl);
        split8(ui, ih, il);
        u4 *d = reinterpret_cast<u4 *>(dst);
        d[0 * 64 + l] = rh;
        d[1 * 64 + l] = rl;
        d[2 * 64 + l] = ih;
        d[3 * 64 + l] = il;
    } else {
        f4 *d = reinterpret_cast<f4 *>(dst);
        d[(0 * 64 + l) * 2 + 0] = f4{ur[0], ur[1], ur[2], ur[3]};
        d[(0 * 64 + l) * 2 + 1] = f4{ur[4], ur[5], ur[6], ur[7]};
        d[(1 * 64 + l) * 2 + 0] = f4{ui[0], ui[1], ui[2], ui[3]};
        d[(1 * 64 + l) * 2 + 1] = f4{ui[4], ui[5], ui[6], ui[7]};
    }
}

constexpr int kMfNS = 5;     // max 8-slot groups per thread in the steering generation
constexpr int kMfWaves = 8;  // MFMA waves of a workgroup (both kernels)
constexpr int kGenWaves = 4; // steering waves of the warp-specialised kernel

// float64 distances from grid point (gx, gy, gz) to the microphones of this thread's 8-slot
// groups gi = part + TPC s (group gi: block gi / 4, lane group gi % 4) into rr; bit s*8+e of
// live marks a slot that holds a real microphone of a real grid point.  Returns this thread's
// share of the steering normalisation sum (sum r^2 or sum 1/r^2).
__device__ __forceinline__ double steer_distances(const MfJob &jb, bool valid, double gx, double gy, double gz,
                                                  int part, int TPC, int ngroups, double (&rr)[kMfNS][8],
                                                  uint32_t &live)
{
    const int M = jb.M;
    const ABF_GLOBAL double *mpos = gp(jb.mpos);
    double help = 0.0;
    live = 0;
#pragma unroll
    for (int s = 0; s < kMfNS; ++s) {
        const int gi = part + TPC * s;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = (gi >> 2) * 32 + kslot_mic(gi & 3, e);
            double r = 1.0;
            if (gi < ngroups && valid && m < M) {
                r = dist3(gx, gy, gz, mpos[m], mpos[M + m], mpos[2 * M + m]);
                help += help_term(jb.steer_type, r);
                live |= 1u << (s * 8 + e);
            }
            rr[s][e] = r;
        }
    }
    return help;
}

// Steering matrix X of one frequency (wavenumber k): this thread's slot groups of tile column
// col into the LDS planes at xbuf (bf16x3: re hi, re lo, im hi, im lo; fp32: re, im; [BN][KS]
// each, plane stride pe elements).  One straight-line body per trig variant, so the 8
// independent phasors of a group interleave; steer and phase variants are selects.
template <bool BF16X3, bool FAST>
__device__ __forceinline__ void steer_generate(const double (&rr)[kMfNS][8], uint32_t live, double k, bool refph,
                                               int wk, int col, int part, int TPC, int ngroups, int KS, int64_t pe,
                                               char *xbuf)
{
    uint16_t *xh16 = reinterpret_cast<uint16_t *>(xbuf);
    float *x32 = reinterpret_cast<float *>(xbuf);
#pragma unroll
    for (int s = 0; s < kMfNS; ++s) {
        const int gi = part + TPC * s;
        if (gi < ngroups) {
            float xr[8], xi[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const double r = rr[s][e];
                double phi = k * r;
                phi = refph ? (double)(float)phi : phi;
                float sn, c;
                if constexpr (FAST) {
                    // hardware sin/cos take revolutions; reduce to [-1/2, 1/2] in float64
                    const double u = phi * kInvTwoPi;
                    const float t = (float)(u - rint(u));
                    sn = __builtin_amdgcn_sinf(t);
                    c = __builtin_amdgcn_cosf(t);
                } else {
                    const double n = rint(phi * kInvTwoPi);
                    double t = fma(-n, kTwoPiHi, phi);
                    t = fma(-n, kTwoPiLo, t);
                    sincosf((float)t, &sn, &c);
                }
                const float rf = (float)r;
                float w = wk == 0 ? 1.f : (wk == 1 ? rf : __builtin_amdgcn_rcpf(rf));
                w = (live >> (s * 8 + e)) & 1u ? w : 0.f;
                xr[e] = w * c;
                xi[e] = -w * sn;
            }
            const int64_t off = (int64_t)col * KS + gi * 8;
            if constexpr (BF16X3) {
                u4 rh, rl, ih, il;
                split8(xr, rh, rl);
                split8(xi, ih, il);
                *reinterpret_cast<u4 *>(xh16 + off) = rh;
                *reinterpret_cast<u4 *>(xh16 + pe + off) = rl;
                *reinterpret_cast<u4 *>(xh16 + 2 * pe + off) = ih;
                *reinterpret_cast<u4 *>(xh16 + 3 * pe + off) = il;
            } else {
                *reinterpret_cast<f4 *>(x32 + off) = f4{xr[0], xr[1], xr[2], xr[3]};
                *reinterpret_cast<f4 *>(x32 + off + 4) = f4{xr[4], xr[5], xr[6], xr[7]};
                *reinterpret_cast<f4 *>(x32 + pe + off) = f4{xi[0], xi[1], xi[2], xi[3]};
                *reinterpret_cast<f4 *>(x32 + pe + off + 4) = f4{xi[4], xi[5], xi[6], xi[7]};
            }
        }
    }
}

template <bool BF16X3>
__device__ __forceinline__ void steer_generate_any(bool fast, const double (&rr)[kMfNS][8], uint32_t live, double k,
                                                   bool refph, int wk, int col, int part, int TPC, int ngroups,
                                                   int KS, int64_t pe, char *xbuf)
{
    if (fast)
        steer_generate<BF16X3, true>(rr, live, k, refph, wk, col, part, TPC, ngroups, KS, pe, xbuf);
    else
        steer_generate<BF16X3, false>(rr, live, k, refph, wk, col, part, TPC, ngroups, KS, pe, xbuf);
}

// Y = U X over the upper block triangle of one frequency and q = Re sum conj(x) Y, by the
// kMfWaves MFMA waves: each takes (row tile, column group) units in a boustrophedon order that
// balances the block triangle, streams its U fragments from L2 (prefetched one block ahead,
// each reused across CW column tiles) and accumulates Yr, Yi in MFMA registers.  The partial q
// of each wave and tile column is added into s_q[wave][BN].
template <bool BF16X3, int CW>
__device__ __forceinline__ void gemm_frequency(int wave, int lane, int nK, int NCG, const ABF_GLOBAL char *prep_f,
                                               const char *xbuf, int64_t pe, int KS, int BN, float *s_q)
{
    const uint16_t *xh16 = reinterpret_cast<const uint16_t *>(xbuf);
    const float *x32 = reinterpret_cast<const float *>(xbuf);
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    const int nunits = 2 * nK * NCG;
    const int rounds = (nunits + kMfWaves - 1) / kMfWaves;
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
        const ABF_GLOBAL char *ablk = prep_f + ((int64_t)upper_index(I, I, nK) * 2 + t) * 4096;
        const int colbase = cg * CW * 16 + (lane & 15);
        const int kq = (lane >> 4) * 8;
        if constexpr (BF16X3) {
            const ABF_GLOBAL u4 *ap = reinterpret_cast<const ABF_GLOBAL u4 *>(ablk);
            u4 a_rh = ap[lane], a_rl = ap[64 + lane], a_ih = ap[128 + lane], a_il = ap[192 + lane];
            for (int J = I; J < nK; ++J) {
                // prefetch the next block's fragments (blocks of a row are 8 KiB apart)
                u4 n_rh = a_rh, n_rl = a_rl, n_ih = a_ih, n_il = a_il;
                if (J + 1 < nK) {
                    const ABF_GLOBAL u4 *np = reinterpret_cast<const ABF_GLOBAL u4 *>(ablk + (int64_t)(J + 1 - I) * 8192);
                    n_rh = np[lane];
                    n_rl = np[64 + lane];
                    n_ih = np[128 + lane];
                    n_il = np[192 + lane];
                }
                const u4 a_nh = a_ih ^ sg, a_nl = a_il ^ sg;
                const int koff = J * 32 + kq;
#define MF(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A), B, C, 0, 0, 0)
                bf16x8 Xrh[CW], Xrl[CW], Xih[CW], Xil[CW];
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int64_t xo = (int64_t)(colbase + c * 16) * KS + koff;
                    Xrh[c] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + xo));
                    Xrl[c] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + pe + xo));
                    Xih[c] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + 2 * pe + xo));
                    Xil[c] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u4 *>(xh16 + 3 * pe + xo));
                }
                // Yr = Ur Xr - Ui Xi ; Yi = Ui Xr + Ur Xi, each product hi*hi + hi*lo + lo*hi;
                // 2 CW independent accumulators between dependent MFMAs
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rh, Xrh[c], accr[c]); MF(a_ih, Xrh[c], acci[c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rh, Xrl[c], accr[c]); MF(a_ih, Xrl[c], acci[c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rl, Xrh[c], accr[c]); MF(a_il, Xrh[c], acci[c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nh, Xih[c], accr[c]); MF(a_rh, Xih[c], acci[c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nh, Xil[c], accr[c]); MF(a_rh, Xil[c], acci[c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nl, Xih[c], accr[c]); MF(a_rl, Xih[c], acci[c]); }
#undef MF
                a_rh = n_rh;
                a_rl = n_rl;
                a_ih = n_ih;
                a_il = n_il;
            }
        } else {
            const ABF_GLOBAL f4 *ap = reinterpret_cast<const ABF_GLOBAL f4 *>(ablk);
            f4 u_a = ap[2 * lane], u_b = ap[2 * lane + 1], v_a = ap[128 + 2 * lane], v_b = ap[128 + 2 * lane + 1];
            for (int J = I; J < nK; ++J) {
                f4 nu_a = u_a, nu_b = u_b, nv_a = v_a, nv_b = v_b;
                if (J + 1 < nK) {
                    const ABF_GLOBAL f4 *np = reinterpret_cast<const ABF_GLOBAL f4 *>(ablk + (int64_t)(J + 1 - I) * 8192);
                    nu_a = np[2 * lane];
                    nu_b = np[2 * lane + 1];
                    nv_a = np[128 + 2 * lane];
                    nv_b = np[128 + 2 * lane + 1];
                }
                const int koff = J * 32 + kq;
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int64_t xo = (int64_t)(colbase + c * 16) * KS + koff;
                    const f4 xr0 = *reinterpret_cast<const f4 *>(x32 + xo);
                    const f4 xr1 = *reinterpret_cast<const f4 *>(x32 + xo + 4);
                    const f4 xi0 = *reinterpret_cast<const f4 *>(x32 + pe + xo);
                    const f4 xi1 = *reinterpret_cast<const f4 *>(x32 + pe + xo + 4);
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float ur = e < 4 ? u_a[e & 3] : u_b[e & 3], ui = e < 4 ? v_a[e & 3] : v_b[e & 3];
                        const float xr = e < 4 ? xr0[e & 3] : xr1[e & 3], xi = e < 4 ? xi0[e & 3] : xi1[e & 3];
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xr, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui, xr, acci[c], 0, 0, 0);
                        accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui, xi, accr[c], 0, 0, 0);
                        acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xi, acci[c], 0, 0, 0);
                    }
                }
                u_a = nu_a;
                u_b = nu_b;
                v_a = nv_a;
                v_b = nv_b;
            }
        }
        // epilogue: result rows 4 g + i of row tile t sit at K-slot 8 g + 4 t + i of block I
        const int slot0 = I * 32 + (lane >> 4) * 8 + t * 4;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int colc = colbase + c * 16;
            const int64_t xo = (int64_t)colc * KS + slot0;
            float sacc = 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float xr, xi;
                if constexpr (BF16X3) {
                    xr = bf16_to_float(xh16[xo + e]) + bf16_to_float(xh16[pe + xo + e]);
                    xi = bf16_to_float(xh16[2 * pe + xo + e]) + bf16_to_float(xh16[3 * pe + xo + e]);
                } else {
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
}

// Normalise and store frequency f of the tile: q = sum of the NQ partials in s_q[NQ][BN]
// (fixed order, then cleared), steering sum = sum of help_parts partials of s_help[part][BN].
template <int NQ>
__device__ __forceinline__ void store_frequency(const MfJob &jb, int f, int tile, int BN, int idx0, int stride,
                                                float *s_q, const double *s_r0, const double *s_help,
                                                int help_parts)
{
    const int64_t G = jb.G;
    for (int cidx = idx0; cidx < BN; cidx += stride) {
        double qs = 0.0;
#pragma unroll
        for (int w = 0; w < NQ; ++w) {
            qs += (double)s_q[w * BN + cidx];
            s_q[w * BN + cidx] = 0.f;
        }
        double help = 0.0;
        for (int p = 0; p < help_parts; ++p) help += s_help[p * BN + cidx];
        const int64_t gg = (int64_t)tile * BN + cidx;
        if (gg < G) {
            double res, nst;
            normalise(jb.steer_type, qs, s_r0[cidx], help, (double)jb.M, jb.norm, res, nst);
            float rf = (float)res;
            if (jb.clip && rf < 0.f) rf = 0.f;
            gpw(jb.out)[(int64_t)f * G + gg] = rf;
            if (jb.out_norm) gpw(jb.out_norm)[(int64_t)f * G + gg] = (float)nst;
        }
    }
}

// One workgroup = 8 waves computes a tile of BN = 16 CT grid points against all Mp microphones
// for FB frequencies.  Prologue: float64 distances of the workgroup's (point, microphone) pairs
// into registers, steering-normalisation sums and r0.  Per frequency: all waves generate the
// steering matrix X of the tile (BN x Mp, split hi+lo bf16 or fp32, K-slot order of kslot_mic)
// into LDS, then run gemm_frequency; q is reduced per grid point in a fixed order
// (bit-identical results alone and batched), then normalised per steer type.
template <bool BF16X3, int CW>
__global__ __launch_bounds__(kMfWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void bf_mfma_kernel(
    const MfJob *__restrict__ jobs, int njobs)
{
    constexpr int NT = kMfWaves * 64;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t wg = blockIdx.x;
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];

    const int Mp = jb.Mp, CT = jb.CT, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int BN = 16 * CT;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);
    const int KS = BF16X3 ? Mp + 8 : Mp + 4; // LDS row stride in elements (16-byte skew)
    const int nK = Mp / 32;
    const int nUp = nK * (nK + 1) / 2;

    // LDS: X planes [BN][KS] (8 bytes per element in both forms), then r0, help, q
    const int64_t pe = (int64_t)BN * KS; // plane stride in elements
    double *s_r0 = reinterpret_cast<double *>(smem + pe * 8);
    double *s_help = s_r0 + BN;
    float *s_q = reinterpret_cast<float *>(s_help + BN); // [kMfWaves][BN]

    // ---- prologue: distances of this thread's (point, 8-slot groups) into registers ----
    const int TPC = NT / BN; // threads per grid point (power of two)
    const int col = tid / TPC, part = tid % TPC;
    const int64_t g = (int64_t)tile * BN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = gp(jb.gpos)[g];
        gy = gp(jb.gpos)[G + g];
        gz = gp(jb.gpos)[2 * G + g];
    }
    const int ngroups = Mp / 8;
    double rr[kMfNS][8];
    uint32_t live;
    double help = steer_distances(jb, valid, gx, gy, gz, part, TPC, ngroups, rr, live);
    for (int off = TPC >> 1; off > 0; off >>= 1) help += __shfl_xor(help, off);
    if (part == 0) {
        s_help[col] = help;
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
        s_r0[col] = r0;
    }
    for (int idx = tid; idx < kMfWaves * BN; idx += NT) s_q[idx] = 0.f;

    const int wk = weight_kind(jb.steer_type);
    const bool refph = jb.phase_mode == ABF_PHASE_REFERENCE;
    const bool fast = jb.fast_trig != 0;
    const int NCG = CT / CW;
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));

    // Frequencies in band order: workgroups running side by side on one XCD then read the U
    // blocks of the same one or two frequencies at any moment, which keeps them in that XCD's L2.
    for (int f = f0; f < f1; ++f) {
        steer_generate_any<BF16X3>(fast, rr, live, gp(jb.wavenumber)[f], refph, wk, col, part, TPC, ngroups, KS, pe,
                                   smem);
        __syncthreads();
        gemm_frequency<BF16X3, CW>(wave, lane, nK, NCG, prep_base + (int64_t)f * nUp * 8192, smem, pe, KS, BN, s_q);
        __syncthreads();
        store_frequency<kMfWaves>(jb, f, tile, BN, tid, NT, s_q, s_r0, s_help, 1);
        // the next frequency overwrites X only after every wave has passed this barrier
        __syncthreads();
    }
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads (the U fragment
// prefetches of the MFMA waves) stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kWsMf = 4; // MFMA waves of the warp-specialised kernel (one per SIMD)
constexpr int kWsCW = 2; // column tiles per U fragment in the warp-specialised kernel

// Warp-specialised form: 4 MFMA waves and 4 steering waves (one of each per SIMD), X
// double-buffered in LDS.  While the MFMA waves work on X of frequency f, the steering waves
// generate X of frequency f + 1 into the other buffer and normalise/store frequency f - 1, so
// the VALU-heavy generation runs beside the MFMA work instead of alternating with it; one LDS
// barrier per frequency hands the buffers over.  Each MFMA wave walks its (unit, block) steps
// of all FB frequencies as one flat sequence: U fragments are prefetched three steps ahead
// (across unit and frequency boundaries) into a ring of four register sets and X fragments one
// step ahead, so one wave per SIMD keeps its MFMA pipe fed.  The steering waves hold the tile's
// float64 distances in registers for all FB frequencies.
template <bool BF16X3>
__global__ __launch_bounds__((kWsMf + kGenWaves) * 64) void bf_mfma_ws_kernel(const MfJob *__restrict__ jobs,
                                                                              int njobs)
{
    constexpr int NG = kGenWaves * 64;
    constexpr int CW = kWsCW;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t wg = blockIdx.x;
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];

    const int Mp = jb.Mp, CT = jb.CT, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int BN = 16 * CT;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);
    const int nf = f1 - f0;
    const int KS = BF16X3 ? Mp + 8 : Mp + 4;
    const int nK = Mp / 32;
    const int nUp = nK * (nK + 1) / 2;
    // frequencies in band order (see bf_mfma_kernel)
    auto phys = [&](int f) __attribute__((always_inline)) { return f; };

    // LDS: two X buffers [planes][BN][KS] (8 bytes per element), q partials [2][kWsMf][BN],
    // r0 [BN], steering-sum partials [TPC][BN]
    const int64_t pe = (int64_t)BN * KS;
    auto xb = [&](int b) __attribute__((always_inline)) { return smem + b * pe * 8; }; // X buffer b (LDS)
    float *s_q = reinterpret_cast<float *>(smem + pe * 16);
    double *s_r0 = reinterpret_cast<double *>(s_q + 2 * kWsMf * BN);
    double *s_help = s_r0 + BN;

    if (wave >= kWsMf) {
        // ---------------- steering waves ----------------
        const int gt = tid - kWsMf * 64;
        const int TPC = NG / BN;
        const int col = gt % BN, part = gt / BN; // consecutive lanes take consecutive grid points
        const int64_t g = (int64_t)tile * BN + col;
        const bool valid = g < G;
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (valid) {
            gx = gp(jb.gpos)[g];
            gy = gp(jb.gpos)[G + g];
            gz = gp(jb.gpos)[2 * G + g];
        }
        const int ngroups = Mp / 8;
        double rr[kMfNS][8];
        uint32_t live;
        s_help[part * BN + col] = steer_distances(jb, valid, gx, gy, gz, part, TPC, ngroups, rr, live);
        if (part == 0) {
            double r0 = jb.ref_dist;
            if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
            s_r0[col] = r0;
        }
        for (int idx = gt; idx < 2 * kWsMf * BN; idx += NG) s_q[idx] = 0.f;
        const int wk = weight_kind(jb.steer_type);
        const bool refph = jb.phase_mode == ABF_PHASE_REFERENCE;
        const bool fast = jb.fast_trig != 0;
        steer_generate_any<BF16X3>(fast, rr, live, gp(jb.wavenumber)[phys(f0)], refph, wk, col, part, TPC, ngroups, KS, pe,
                                   xb(0));
        lds_barrier();
        for (int f = f0; f < f1; ++f) {
            const int b = (f - f0) & 1;
            if (f + 1 < f1)
                steer_generate_any<BF16X3>(fast, rr, live, gp(jb.wavenumber)[phys(f + 1)], refph, wk, col, part, TPC, ngroups,
                                           KS, pe, xb(b ^ 1));
            if (f > f0)
                store_frequency<kWsMf>(jb, phys(f - 1), tile, BN, gt, NG, s_q + (b ^ 1) * kWsMf * BN, s_r0, s_help, TPC);
            lds_barrier();
        }
        store_frequency<kWsMf>(jb, phys(f1 - 1), tile, BN, gt, NG, s_q + ((nf - 1) & 1) * kWsMf * BN, s_r0, s_help, TPC);
        return;
    }

    // ---------------- MFMA waves ----------------
    // Step table of this wave for one frequency (the same for every frequency): lane s < S
    // holds step s -- its U fragment offset within a frequency, its X fragment offset and, for
    // the last step of a unit, the unit's epilogue slots.  Units u = 4 rd + (rd odd ? 3 - wave :
    // wave) in rounds rd, blocks J = I .. nK - 1 within a unit.
    const int lncg = CT / CW == 2 ? 1 : 0; // column groups per tile: 1 or 2
    const int nunits = 2 * nK << lncg;
    const int rounds = (nunits + kWsMf - 1) / kWsMf;
    int S = 0;
    uint32_t tab_a = 0, tab_x = 0, tab_e = 0;
    for (int rd = 0; rd < rounds; ++rd) {
        const int u = rd * kWsMf + ((rd & 1) ? (kWsMf - 1 - wave) : wave);
        if (u >= nunits) continue;
        const int rt = u >> lncg, cg = u & ((1 << lncg) - 1);
        const int I = rt >> 1, t = rt & 1;
        for (int J = I; J < nK; ++J, ++S) {
            if (lane == S) {
                tab_a = (uint32_t)(upper_index(I, J, nK) * 2 + t) * 4096u;
                tab_x = (uint32_t)(cg * CW * 16 * KS + J * 32);
                tab_e = (J == nK - 1 ? 0x80000000u : 0u) | ((uint32_t)cg << 16) | (uint32_t)(I * 32 + t * 4);
            }
        }
    }
    lds_barrier();
    if (S == 0) {
        for (int f = f0; f < f1; ++f) lds_barrier();
        return;
    }
    const int total = S * nf;
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));
    const int64_t fstride = (int64_t)nUp * 8192; // bytes of U per frequency
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    int lx[CW]; // this lane's element offset inside a step's X fragments
#pragma unroll
    for (int c = 0; c < CW; ++c) lx[c] = (c * 16 + (lane & 15)) * KS + (lane >> 4) * 8;

    // (table index, logical frequency) of steps sk, sk + 1 and sk + 3
    int i0 = 0, g0 = 0, i1 = 0, g1 = 0, i3 = 0, g3 = 0;
    auto adv = [&](int &i, int &g) __attribute__((always_inline)) {
        if (++i == S) {
            i = 0;
            ++g;
        }
    };
    adv(i1, g1);
    adv(i3, g3);
    adv(i3, g3);
    adv(i3, g3);
    auto a_ptr = [&](int i, int g) __attribute__((always_inline)) {
        return prep_base + phys(f0 + g) * fstride + __builtin_amdgcn_readlane(tab_a, i);
    };
    // element offset of buffer (g & 1): 8 bytes per element and buffer -> 4 bf16 or 2 floats
    const int64_t buf_el = BF16X3 ? 4 * pe : 2 * pe;

    if constexpr (BF16X3) {
        const uint16_t *xh = reinterpret_cast<const uint16_t *>(smem);
        u4 a[4][4];     // U fragment ring: re hi, re lo, im hi, im lo
        u4 x[2][CW][4]; // X fragments of the current and the next step
        auto load_a = [&](const ABF_GLOBAL char *p, u4 (&d)[4]) __attribute__((always_inline)) {
            const ABF_GLOBAL u4 *q = reinterpret_cast<const ABF_GLOBAL u4 *>(p) + lane;
#pragma unroll
            for (int j = 0; j < 4; ++j) d[j] = q[64 * j];
        };
        auto load_x = [&](int i, int g, u4 (&d)[CW][4]) __attribute__((always_inline)) {
            const int64_t base = (int64_t)__builtin_amdgcn_readlane(tab_x, i) + (g & 1) * buf_el;
#pragma unroll
            for (int c = 0; c < CW; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) d[c][q] = *reinterpret_cast<const u4 *>(xh + base + q * pe + lx[c]);
        };
        // prologue prefetch in step order, as the loop's wait counts assume
        load_a(a_ptr(0, 0), a[0]);
        __builtin_amdgcn_sched_barrier(0);
        load_a(total > 1 ? a_ptr(i1, g1) : a_ptr(0, 0), a[1]);
        __builtin_amdgcn_sched_barrier(0);
        {
            int i2 = i1, g2 = g1;
            adv(i2, g2);
            load_a(total > 2 ? a_ptr(i2, g2) : a_ptr(0, 0), a[2]);
        }
        __builtin_amdgcn_sched_barrier(0);
        load_x(0, 0, x[0]);
        f4 accr[CW], acci[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            accr[c] = f4{0.f, 0.f, 0.f, 0.f};
            acci[c] = f4{0.f, 0.f, 0.f, 0.f};
        }
        // Every substep issues its prefetches (past the end from step 0's addresses), so that all
        // paths through the loop issue the same loads and the wait counts stay exact.
        auto substep = [&](auto kc, int sk) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            const bool active = sk < total;
            load_a(sk + 3 < total ? a_ptr(i3, g3) : a_ptr(0, 0), a[(k + 3) & 3]);
            adv(i3, g3);
            const bool fend = sk + 1 >= total || i1 == 0;
            if (fend)
                load_x(active ? i0 : 0, active ? g0 : 0, x[(k + 1) & 1]);
            else
                load_x(i1, g1, x[(k + 1) & 1]);
            // keep the prefetches ahead of this step's MFMAs (the scheduler would otherwise sink
            // them below the MFMAs to save registers, which defeats the prefetch)
            __builtin_amdgcn_sched_barrier(0);
            if (!active) return;
            const u4 a_rh = a[k & 3][0], a_rl = a[k & 3][1], a_ih = a[k & 3][2], a_il = a[k & 3][3];
            const u4 a_nh = a_ih ^ sg, a_nl = a_il ^ sg;
#define MF(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A), \
                                                                __builtin_bit_cast(bf16x8, B), C, 0, 0, 0)
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_rh, x[k & 1][c][0], accr[c]); MF(a_ih, x[k & 1][c][0], acci[c]); }
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_rh, x[k & 1][c][1], accr[c]); MF(a_ih, x[k & 1][c][1], acci[c]); }
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_rl, x[k & 1][c][0], accr[c]); MF(a_il, x[k & 1][c][0], acci[c]); }
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_nh, x[k & 1][c][2], accr[c]); MF(a_rh, x[k & 1][c][2], acci[c]); }
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_nh, x[k & 1][c][3], accr[c]); MF(a_rh, x[k & 1][c][3], acci[c]); }
#pragma unroll
            for (int c = 0; c < CW; ++c) { MF(a_nl, x[k & 1][c][2], accr[c]); MF(a_rl, x[k & 1][c][2], acci[c]); }
#undef MF
            const uint32_t e = __builtin_amdgcn_readlane(tab_e, i0);
            if (e >> 31) {
                // unit done: result rows 4 g + i of row tile t sit at K-slot 8 g + 4 t + i of block I
                const int64_t base = (g0 & 1) * buf_el;
                float *sq = s_q + ((g0 & 1) * kWsMf + wave) * BN;
                const int slot0 = (int)(e & 0xffffu) + (lane >> 4) * 8;
                const int cgq = (int)((e >> 16) & 0x7fffu);
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int colc = cgq * CW * 16 + c * 16 + (lane & 15);
                    const int64_t xo = base + (int64_t)colc * KS + slot0;
                    float sacc = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float xr = bf16_to_float(xh[xo + q]) + bf16_to_float(xh[pe + xo + q]);
                        const float xi = bf16_to_float(xh[2 * pe + xo + q]) + bf16_to_float(xh[3 * pe + xo + q]);
                        sacc = fmaf(xr, accr[c][q], sacc);
                        sacc = fmaf(xi, acci[c][q], sacc);
                    }
                    sacc += __shfl_xor(sacc, 16);
                    sacc += __shfl_xor(sacc, 32);
                    if (lane < 16) sq[colc] += sacc;
                    accr[c] = f4{0.f, 0.f, 0.f, 0.f};
                    acci[c] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
            if (fend) {
                lds_barrier();
                if (sk + 1 < total) load_x(i1, g1, x[(k + 1) & 1]);
            }
            i0 = i1;
            g0 = g1;
            adv(i1, g1);
        };
        for (int s0 = 0; s0 < total; s0 += 4) {
            substep(std::integral_constant<int, 0>{}, s0);
            substep(std::integral_constant<int, 1>{}, s0 + 1);
            substep(std::integral_constant<int, 2>{}, s0 + 2);
            substep(std::integral_constant<int, 3>{}, s0 + 3);
        }
    } else {
        // fp32 form: 16x16x4 fp32 MFMA, same step sequence and prefetch distances
        const float *x32 = reinterpret_cast<const float *>(smem);
        f4 a[4][4]; // U fragment ring: re e0-3, re e4-7, im e0-3, im e4-7
        f4 x[2][CW][4];
        auto load_a = [&](const ABF_GLOBAL char *p, f4 (&d)[4]) __attribute__((always_inline)) {
            const ABF_GLOBAL f4 *q = reinterpret_cast<const ABF_GLOBAL f4 *>(p);
            d[0] = q[2 * lane];
            d[1] = q[2 * lane + 1];
            d[2] = q[128 + 2 * lane];
            d[3] = q[128 + 2 * lane + 1];
        };
        auto load_x = [&](int i, int g, f4 (&d)[CW][4]) __attribute__((always_inline)) {
            const int64_t base = (int64_t)__builtin_amdgcn_readlane(tab_x, i) + (g & 1) * buf_el;
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int64_t xo = base + lx[c];
                d[c][0] = *reinterpret_cast<const f4 *>(x32 + xo);
                d[c][1] = *reinterpret_cast<const f4 *>(x32 + xo + 4);
                d[c][2] = *reinterpret_cast<const f4 *>(x32 + pe + xo);
                d[c][3] = *reinterpret_cast<const f4 *>(x32 + pe + xo + 4);
            }
        };
        load_a(a_ptr(0, 0), a[0]);
        __builtin_amdgcn_sched_barrier(0);
        load_a(total > 1 ? a_ptr(i1, g1) : a_ptr(0, 0), a[1]);
        __builtin_amdgcn_sched_barrier(0);
        {
            int i2 = i1, g2 = g1;
            adv(i2, g2);
            load_a(total > 2 ? a_ptr(i2, g2) : a_ptr(0, 0), a[2]);
        }
        __builtin_amdgcn_sched_barrier(0);
        load_x(0, 0, x[0]);
        f4 accr[CW], acci[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            accr[c] = f4{0.f, 0.f, 0.f, 0.f};
            acci[c] = f4{0.f, 0.f, 0.f, 0.f};
        }
        auto substep = [&](auto kc, int sk) __attribute__((always_inline)) {
            constexpr int k = decltype(kc)::value;
            const bool active = sk < total;
            load_a(sk + 3 < total ? a_ptr(i3, g3) : a_ptr(0, 0), a[(k + 3) & 3]);
            adv(i3, g3);
            const bool fend = sk + 1 >= total || i1 == 0;
            if (fend)
                load_x(active ? i0 : 0, active ? g0 : 0, x[(k + 1) & 1]);
            else
                load_x(i1, g1, x[(k + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
            if (!active) return;
#pragma unroll
            for (int c = 0; c < CW; ++c) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float ur = a[k & 3][q >> 2][q & 3], ui = a[k & 3][2 + (q >> 2)][q & 3];
                    const float xr = x[k & 1][c][q >> 2][q & 3], xi = x[k & 1][c][2 + (q >> 2)][q & 3];
                    accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xr, accr[c], 0, 0, 0);
                    acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui, xr, acci[c], 0, 0, 0);
                    accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui, xi, accr[c], 0, 0, 0);
                    acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xi, acci[c], 0, 0, 0);
                }
            }
            const uint32_t e = __builtin_amdgcn_readlane(tab_e, i0);
            if (e >> 31) {
                const int64_t base = (g0 & 1) * buf_el;
                float *sq = s_q + ((g0 & 1) * kWsMf + wave) * BN;
                const int slot0 = (int)(e & 0xffffu) + (lane >> 4) * 8;
                const int cgq = (int)((e >> 16) & 0x7fffu);
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int colc = cgq * CW * 16 + c * 16 + (lane & 15);
                    const int64_t xo = base + (int64_t)colc * KS + slot0;
                    float sacc = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        sacc = fmaf(x32[xo + q], accr[c][q], sacc);
                        sacc = fmaf(x32[pe + xo + q], acci[c][q], sacc);
                    }
                    sacc += __shfl_xor(sacc, 16);
                    sacc += __shfl_xor(sacc, 32);
                    if (lane < 16) sq[colc] += sacc;
                    accr[c] = f4{0.f, 0.f, 0.f, 0.f};
                    acci[c] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
            if (fend) {
                lds_barrier();
                if (sk + 1 < total) load_x(i1, g1, x[(k + 1) & 1]);
            }
            i0 = i1;
            g0 = g1;
            adv(i1, g1);
        };
        for (int s0 = 0; s0 < total; s0 += 4) {
            substep(std::integral_constant<int, 0>{}, s0);
            substep(std::integral_constant<int, 1>{}, s0 + 1);
            substep(std::integral_constant<int, 2>{}, s0 + 2);
            substep(std::integral_constant<int, 3>{}, s0 + 3);
        }
    }
}


template __global__ void bf_mfma_kernel<true, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<true, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<true, 4>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<false, 4>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<false>(const MfJob *, int);

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
