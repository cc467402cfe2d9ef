// MFMA beamformer kernels (device code), included by the translation units that instantiate
// them (abf_bf_mfma_bf16.hip, abf_bf_mfma_f16.hip, abf_bf_mfma_f32.hip).  PR selects the operand
// precision (kPrF32, kPrBF16X3, kPrF16X2).  VAR = fast trig (bit 0) | weight kind << 1
// selects the steering-generation variant at compile time, so each instantiation carries one
// straight-line generation body.
#pragma once
#include <type_traits>

#include "abf_mfma.h"

namespace abf {

constexpr int kMfNS = 5;     // max 8-slot groups per thread in the steering generation
constexpr int kMfWaves = 8;  // MFMA waves of a workgroup (both kernels)
constexpr int kGenWaves = 4; // steering waves of the warp-specialised kernel

// float64 distances from grid point (gx, gy, gz) to the microphones of this thread's 8-slot
// groups gi = part + TPC s (group gi: block gi / 4, lane group gi % 4) into rr (1 for slots
// without a real microphone or grid point).  Returns this thread's share of the steering
// normalisation sum (sum r^2 or sum 1/r^2).
template <int NS>
__device__ __forceinline__ double steer_distances(const MfJob &jb, bool valid, double gx, double gy, double gz,
                                                  int part, int TPC, int ngroups, double (&rr)[NS][8])
{
    const int M = jb.M;
    const ABF_GLOBAL double *mpos = gp(jb.mpos);
    double help = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int gi = part + TPC * s;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = (gi >> 2) * 32 + kslot_mic(gi & 3, e);
            double r = 1.0;
            if (gi < ngroups && valid && m < M) {
                r = dist3(gx, gy, gz, mpos[m], mpos[M + m], mpos[2 * M + m]);
                help += help_term(jb.steer_type, r);
            }
            rr[s][e] = r;
        }
    }
    return help;
}

// Steering matrix X of one frequency (wavenumber k): this thread's slot groups of tile column
// col into the LDS planes at xbuf (bf16x3: re hi, re lo, im hi, im lo; fp32: re, im; [BN][KS]
// each, plane stride pe elements).  One straight-line body per (trig, weight, phase) variant, so
// the 8 independent phasors of a group interleave without selects.  Slots without a real
// microphone or grid point get finite values (r = 1): the re-laid-out CSM is zero in the rows and
// columns of padded microphones and the maps of padded grid points are not stored.
// With PR = kPrF16X2 the weights are multiplied by xs, a power of two that brings them near 1
// (fp16 keeps 5 exponent bits); the epilogue divides q by xs^2.
template <int PR, bool FAST, int WK, int NS>
__device__ __forceinline__ void steer_generate(const double (&rr)[NS][8], double k, bool refph, int col, int part,
                                               int TPC, int ngroups, int KS, int64_t pe, char *xbuf, float xs = 1.f)
{
    constexpr bool BF16X3 = PR == kPrBF16X3;
    uint16_t *xh16 = reinterpret_cast<uint16_t *>(xbuf);
    float *x32 = reinterpret_cast<float *>(xbuf);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int gi = part + TPC * s;
        if (gi < ngroups) {
            float xr[8], xi[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                double r = rr[s][e];
                // opaque to the optimiser: keeps the frequency-invariant weight (float r or 1/r)
                // from being hoisted out of the frequency loop into 40 extra live registers
                asm volatile("" : "+v"(r));
                double phi = k * r;
                phi = refph ? (double)(float)phi : phi; // the reference's float32 phase
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
                if constexpr (WK == 0) {
                    xr[e] = c;
                    xi[e] = -sn;
                } else {
                    const float rf = (float)r;
                    float w = WK == 1 ? rf : __builtin_amdgcn_rcpf(rf);
                    if constexpr (PR == kPrF16X2) w *= xs;
                    xr[e] = w * c;
                    xi[e] = -w * sn;
                }
            }
            const int64_t off = (int64_t)col * KS + gi * 8;
            if constexpr (PR != kPrF32) {
                u4 rh, rl, ih, il;
                if constexpr (BF16X3) {
                    split8(xr, rh, rl);
                    split8(xi, ih, il);
                } else {
                    split8_h(xr, rh, rl);
                    split8_h(xi, ih, il);
                }
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
        // one group of 8 phasors in flight at a time (bounded register pressure)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Y = U X over the upper block triangle of one frequency and q = Re sum conj(x) Y, by the
// kMfWaves MFMA waves: each takes (row tile(s), column group) units in a boustrophedon order that
// balances the block triangle, streams its U fragments from L2 (prefetched one block ahead, each
// reused across CW column tiles) and accumulates Yr, Yi in MFMA registers.  With TT = 2 a unit
// covers both 16-row halves of a 32-row block, so every X fragment read from LDS feeds twice the
// MFMAs (half the LDS traffic per MFMA, twice the U traffic).  The partial q of each wave and tile
// column is added into s_q[wave][BN].
template <int PR, int CW, int TT>
__device__ __forceinline__ void gemm_frequency(int wave, int lane, int nK, int NCG, const ABF_GLOBAL char *prep_f,
                                               const char *xbuf, int64_t pe, int KS, int BN, float *s_q)
{
    constexpr bool BF16X3 = PR == kPrBF16X3;
    const uint16_t *xh16 = reinterpret_cast<const uint16_t *>(xbuf);
    const float *x32 = reinterpret_cast<const float *>(xbuf);
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    const int nunits = (2 / TT) * nK * NCG;
    const int rounds = (nunits + kMfWaves - 1) / kMfWaves;
    for (int rd = 0; rd < rounds; ++rd) {
        const int u = rd * kMfWaves + ((rd & 1) ? (kMfWaves - 1 - wave) : wave);
        if (u >= nunits) continue;
        const int rt = u / NCG, cg = u % NCG;
        const int I = TT == 2 ? rt : rt >> 1;
        const int t0 = TT == 2 ? 0 : (rt & 1);
        f4 accr[TT][CW], acci[TT][CW];
#pragma unroll
        for (int tt = 0; tt < TT; ++tt)
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                accr[tt][c] = f4{0.f, 0.f, 0.f, 0.f};
                acci[tt][c] = f4{0.f, 0.f, 0.f, 0.f};
            }
        // U fragments of block (I, J), row half t: prep_f + ((upper(I, J) * 2 + t) * 4 KiB; the
        // blocks of a row are 8 KiB apart
        const ABF_GLOBAL char *ablk = prep_f + ((int64_t)upper_index(I, I, nK) * 2 + t0) * 4096;
        const int colbase = cg * CW * 16 + (lane & 15);
        const int kq = (lane >> 4) * 8;
        if constexpr (PR == kPrF16X2) {
            // U rounded once to fp16 (planes 0 and 2 of a fragment: re, im), X split hi + lo:
            // Yr = Ur Xr - Ui Xi, Yi = Ui Xr + Ur Xi with X = Xhi + Xlo, 8 MFMAs per (row tile,
            // column tile) and block
            u4 a[TT][2]; // re, im
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    a[tt][q] = reinterpret_cast<const ABF_GLOBAL u4 *>(ablk + tt * 4096)[128 * q + lane];
            for (int J = I; J < nK; ++J) {
                u4 n[TT][2];
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 2; ++q) n[tt][q] = a[tt][q];
                if (J + 1 < nK) {
                    const ABF_GLOBAL char *nb = ablk + (int64_t)(J + 1 - I) * 8192;
#pragma unroll
                    for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            n[tt][q] = reinterpret_cast<const ABF_GLOBAL u4 *>(nb + tt * 4096)[128 * q + lane];
                }
                const int koff = J * 32 + kq;
#define MH(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, A), B, C, 0, 0, 0)
                f16x8 Xrh[CW], Xrl[CW], Xih[CW], Xil[CW];
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const int64_t xo = (int64_t)(colbase + c * 16) * KS + koff;
                    Xrh[c] = __builtin_bit_cast(f16x8, *reinterpret_cast<const u4 *>(xh16 + xo));
                    Xrl[c] = __builtin_bit_cast(f16x8, *reinterpret_cast<const u4 *>(xh16 + pe + xo));
                    Xih[c] = __builtin_bit_cast(f16x8, *reinterpret_cast<const u4 *>(xh16 + 2 * pe + xo));
                    Xil[c] = __builtin_bit_cast(f16x8, *reinterpret_cast<const u4 *>(xh16 + 3 * pe + xo));
                }
#pragma unroll
                for (int tt = 0; tt < TT; ++tt) {
                    const u4 a_r = a[tt][0], a_i = a[tt][1];
                    const u4 a_n = a_i ^ sg;
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MH(a_r, Xrh[c], accr[tt][c]); MH(a_i, Xrh[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MH(a_r, Xrl[c], accr[tt][c]); MH(a_i, Xrl[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MH(a_n, Xih[c], accr[tt][c]); MH(a_r, Xih[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MH(a_n, Xil[c], accr[tt][c]); MH(a_r, Xil[c], acci[tt][c]); }
                }
#undef MH
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 2; ++q) a[tt][q] = n[tt][q];
            }
        } else if constexpr (BF16X3) {
            u4 a[TT][4]; // re hi, re lo, im hi, im lo
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int q = 0; q < 4; ++q) a[tt][q] = reinterpret_cast<const ABF_GLOBAL u4 *>(ablk + tt * 4096)[64 * q + lane];
            for (int J = I; J < nK; ++J) {
                // prefetch the next block's fragments
                u4 n[TT][4];
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) n[tt][q] = a[tt][q];
                if (J + 1 < nK) {
                    const ABF_GLOBAL char *nb = ablk + (int64_t)(J + 1 - I) * 8192;
#pragma unroll
                    for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                        for (int q = 0; q < 4; ++q) n[tt][q] = reinterpret_cast<const ABF_GLOBAL u4 *>(nb + tt * 4096)[64 * q + lane];
                }
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
                // 2 TT CW independent accumulators between dependent MFMAs
#pragma unroll
                for (int tt = 0; tt < TT; ++tt) {
                    const u4 a_rh = a[tt][0], a_rl = a[tt][1], a_ih = a[tt][2], a_il = a[tt][3];
                    const u4 a_nh = a_ih ^ sg, a_nl = a_il ^ sg;
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_rh, Xrh[c], accr[tt][c]); MF(a_ih, Xrh[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_rh, Xrl[c], accr[tt][c]); MF(a_ih, Xrl[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_rl, Xrh[c], accr[tt][c]); MF(a_il, Xrh[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_nh, Xih[c], accr[tt][c]); MF(a_rh, Xih[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_nh, Xil[c], accr[tt][c]); MF(a_rh, Xil[c], acci[tt][c]); }
#pragma unroll
                    for (int c = 0; c < CW; ++c) { MF(a_nl, Xih[c], accr[tt][c]); MF(a_rl, Xih[c], acci[tt][c]); }
                }
#undef MF
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) a[tt][q] = n[tt][q];
            }
        } else {
            f4 a[TT][4]; // re e0-3, re e4-7, im e0-3, im e4-7
            auto load_a = [&](const ABF_GLOBAL char *p, f4 (&d)[4]) __attribute__((always_inline)) {
                const ABF_GLOBAL f4 *q = reinterpret_cast<const ABF_GLOBAL f4 *>(p);
                d[0] = q[2 * lane];
                d[1] = q[2 * lane + 1];
                d[2] = q[128 + 2 * lane];
                d[3] = q[128 + 2 * lane + 1];
            };
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) load_a(ablk + tt * 4096, a[tt]);
            for (int J = I; J < nK; ++J) {
                f4 n[TT][4];
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) n[tt][q] = a[tt][q];
                if (J + 1 < nK) {
#pragma unroll
                    for (int tt = 0; tt < TT; ++tt) load_a(ablk + (int64_t)(J + 1 - I) * 8192 + tt * 4096, n[tt]);
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
                    for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float ur = a[tt][e >> 2][e & 3], ui = a[tt][2 + (e >> 2)][e & 3];
                            const float xr = e < 4 ? xr0[e & 3] : xr1[e & 3], xi = e < 4 ? xi0[e & 3] : xi1[e & 3];
                            accr[tt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xr, accr[tt][c], 0, 0, 0);
                            acci[tt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui, xr, acci[tt][c], 0, 0, 0);
                            accr[tt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui, xi, accr[tt][c], 0, 0, 0);
                            acci[tt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, xi, acci[tt][c], 0, 0, 0);
                        }
                }
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) a[tt][q] = n[tt][q];
            }
        }
        // epilogue: result rows 4 g + i of row tile t sit at K-slot 8 g + 4 t + i of block I
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
            const int slot0 = I * 32 + (lane >> 4) * 8 + (t0 + tt) * 4;
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
                    } else if constexpr (PR == kPrF16X2) {
                        xr = f16_to_float(xh16[xo + e]) + f16_to_float(xh16[pe + xo + e]);
                        xi = f16_to_float(xh16[2 * pe + xo + e]) + f16_to_float(xh16[3 * pe + xo + e]);
                    } else {
                        xr = x32[xo + e];
                        xi = x32[pe + xo + e];
                    }
                    sacc = fmaf(xr, accr[tt][c][e], sacc);
                    sacc = fmaf(xi, acci[tt][c][e], sacc);
                }
                sacc += __shfl_xor(sacc, 16);
                sacc += __shfl_xor(sacc, 32);
                if (lane < 16) s_q[wave * BN + colc] += sacc;
            }
        }
    }
}

// Split-bf16 units with a software pipeline: the X fragments (LDS) and U fragments (L2) of block
// J+1 are requested before the MFMAs of block J, into a second register set, so a wave does not
// wait on LDS at the top of every block.  The J loop is unrolled by two so the two register sets
// alternate without copies; past the last block the loads repeat block J (in bounds) instead of
// branching, which keeps the wait counts exact.  TT row tiles (1 or 2 halves of a 32-row block)
// per unit, CW column tiles per U fragment.
template <int CW, int TT>
__device__ __forceinline__ void gemm_frequency_pf(int wave, int lane, int nK, int NCG, const ABF_GLOBAL char *prep_f,
                                                  const char *xbuf, int64_t pe, int KS, int BN, float *s_q)
{
    const uint16_t *xh16 = reinterpret_cast<const uint16_t *>(xbuf);
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    const int nunits = (2 / TT) * nK * NCG;
    const int rounds = (nunits + kMfWaves - 1) / kMfWaves;
    for (int rd = 0; rd < rounds; ++rd) {
        const int u = rd * kMfWaves + ((rd & 1) ? (kMfWaves - 1 - wave) : wave);
        if (u >= nunits) continue;
        const int rt = u / NCG, cg = u % NCG;
        const int I = TT == 2 ? rt : rt >> 1;
        const int t0 = TT == 2 ? 0 : (rt & 1);
        f4 accr[TT][CW], acci[TT][CW];
#pragma unroll
        for (int tt = 0; tt < TT; ++tt)
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                accr[tt][c] = f4{0.f, 0.f, 0.f, 0.f};
                acci[tt][c] = f4{0.f, 0.f, 0.f, 0.f};
            }
        const ABF_GLOBAL char *ablk = prep_f + ((int64_t)upper_index(I, I, nK) * 2 + t0) * 4096; // (I, I), half t0
        const int colbase = cg * CW * 16 + (lane & 15);
        const int kq = (lane >> 4) * 8;
        auto load_a = [&](int J, u4 (&d)[TT][4]) __attribute__((always_inline)) {
            const ABF_GLOBAL char *b = ablk + (int64_t)(J - I) * 8192;
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int q = 0; q < 4; ++q) d[tt][q] = reinterpret_cast<const ABF_GLOBAL u4 *>(b + tt * 4096)[64 * q + lane];
        };
        auto load_x = [&](int J, u4 (&d)[CW][4]) __attribute__((always_inline)) {
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int64_t xo = (int64_t)(colbase + c * 16) * KS + J * 32 + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q) d[c][q] = *reinterpret_cast<const u4 *>(xh16 + q * pe + xo);
            }
        };
        auto block = [&](const u4 (&a)[TT][4], const u4 (&x)[CW][4]) __attribute__((always_inline)) {
#define MF(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A), \
                                                                __builtin_bit_cast(bf16x8, B), C, 0, 0, 0)
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                const u4 a_rh = a[tt][0], a_rl = a[tt][1], a_ih = a[tt][2], a_il = a[tt][3];
                const u4 a_nh = a_ih ^ sg, a_nl = a_il ^ sg;
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rh, x[c][0], accr[tt][c]); MF(a_ih, x[c][0], acci[tt][c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rh, x[c][1], accr[tt][c]); MF(a_ih, x[c][1], acci[tt][c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_rl, x[c][0], accr[tt][c]); MF(a_il, x[c][0], acci[tt][c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nh, x[c][2], accr[tt][c]); MF(a_rh, x[c][2], acci[tt][c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nh, x[c][3], accr[tt][c]); MF(a_rh, x[c][3], acci[tt][c]); }
#pragma unroll
                for (int c = 0; c < CW; ++c) { MF(a_nl, x[c][2], accr[tt][c]); MF(a_rl, x[c][2], acci[tt][c]); }
            }
#undef MF
        };
        u4 a0[TT][4], a1[TT][4], x0[CW][4], x1[CW][4];
        load_a(I, a0);
        load_x(I, x0);
        for (int J = I;;) {
            const int Jn = J + 1 < nK ? J + 1 : J;
            load_a(Jn, a1);
            load_x(Jn, x1);
            __builtin_amdgcn_sched_barrier(0);
            block(a0, x0);
            if (++J >= nK) break;
            const int Jm = J + 1 < nK ? J + 1 : J;
            load_a(Jm, a0);
            load_x(Jm, x0);
            __builtin_amdgcn_sched_barrier(0);
            block(a1, x1);
            if (++J >= nK) break;
        }
        // epilogue: result rows 4 g + i of row half t sit at K-slot 8 g + 4 t + i of block I
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
            const int slot0 = I * 32 + (lane >> 4) * 8 + (t0 + tt) * 4;
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const int colc = colbase + c * 16;
                const int64_t xo = (int64_t)colc * KS + slot0;
                float sacc = 0.f;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float xr = bf16_to_float(xh16[xo + e]) + bf16_to_float(xh16[pe + xo + e]);
                    const float xi = bf16_to_float(xh16[2 * pe + xo + e]) + bf16_to_float(xh16[3 * pe + xo + e]);
                    sacc = fmaf(xr, accr[tt][c][e], sacc);
                    sacc = fmaf(xi, acci[tt][c][e], sacc);
                }
                sacc += __shfl_xor(sacc, 16);
                sacc += __shfl_xor(sacc, 32);
                if (lane < 16) s_q[wave * BN + colc] += sacc;
            }
        }
    }
}

// Normalise and store frequency f of the tile: q = sum of the NQ partials in s_q[NQ][BN]
// (fixed order, then cleared), steering sum = sum of help_parts partials of s_help[part][BN].
// With the f16x2 form, q is multiplied by the powers of two that undo the scaling of U (fsc) and
// of the grid point's steering weights (s_xf[col]); both are exact.
template <int NQ>
__device__ __forceinline__ void store_frequency(const MfJob &jb, int f, int tile, int BN, int idx0, int stride,
                                                float *s_q, const double *s_r0, const double *s_help,
                                                int help_parts, double fsc = 1.0, const double *s_xf = nullptr)
{
    const int64_t G = jb.G;
    for (int cidx = idx0; cidx < BN; cidx += stride) {
        double qs = 0.0;
#pragma unroll
        for (int w = 0; w < NQ; ++w) {
            qs += (double)s_q[w * BN + cidx];
            s_q[w * BN + cidx] = 0.f;
        }
        if (s_xf) qs = qs * fsc * s_xf[cidx];
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
template <int PR, int CW, int VAR, int TT, bool PF>
__global__ __launch_bounds__(kMfWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void bf_mfma_kernel(
    const MfJob *__restrict__ jobs, int njobs)
{
    constexpr bool BF16X3 = PR == kPrBF16X3;
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
    const int KS = PR != kPrF32 ? Mp + 8 : Mp + 4; // LDS row stride in elements (16-byte skew)
    const int nK = Mp / 32;
    const int nUp = nK * (nK + 1) / 2;

    // LDS: X planes [BN][KS] (8 bytes per element in all forms), then r0, help, q, weight scales
    const int64_t pe = (int64_t)BN * KS; // plane stride in elements
    double *s_r0 = reinterpret_cast<double *>(smem + pe * 8);
    double *s_help = s_r0 + BN;
    float *s_q = reinterpret_cast<float *>(s_help + BN); // [kMfWaves][BN]
    double *s_xf = reinterpret_cast<double *>(s_q + kMfWaves * BN); // [BN], f16x2 form only

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
    constexpr int NS = PF ? 4 : kMfNS; // pipelined kernels keep registers for the X prefetch
    double rr[NS][8];
    double help = steer_distances<NS>(jb, valid, gx, gy, gz, part, TPC, ngroups, rr);
    for (int off = TPC >> 1; off > 0; off >>= 1) help += __shfl_xor(help, off);
    // f16x2: the steering weights (r or 1/r) of this grid point are scaled by xs = 2^-e, e the
    // exponent of its weight at microphone 0, so they stay well inside fp16's range
    float xs = 1.f;
    int xe = 0;
    if constexpr (PR == kPrF16X2 && (VAR >> 1) != 0) {
        const ABF_GLOBAL double *mp = gp(jb.mpos);
        const double r = valid ? dist3(gx, gy, gz, mp[0], mp[jb.M], mp[2 * jb.M]) : 1.0;
        const double w0 = (VAR >> 1) == 1 ? r : 1.0 / r;
        xe = (w0 > 0.0 && w0 < INFINITY) ? ilogb(w0) : 0;
        xe = xe < -60 ? -60 : (xe > 60 ? 60 : xe);
        xs = ldexpf(1.f, -xe);
    }
    if (part == 0) {
        s_help[col] = help;
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
        s_r0[col] = r0;
        if constexpr (PR == kPrF16X2) s_xf[col] = ldexp(1.0, 2 * xe);
    }
    for (int idx = tid; idx < kMfWaves * BN; idx += NT) s_q[idx] = 0.f;

    const bool refph = jb.phase_mode == ABF_PHASE_REFERENCE;
    const int NCG = CT / CW;
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));

    // Frequencies in band order: workgroups running side by side on one XCD then read the U
    // blocks of the same one or two frequencies at any moment, which keeps them in that XCD's L2.
    for (int f = f0; f < f1; ++f) {
        steer_generate<PR, (VAR & 1) != 0, (VAR >> 1), NS>(rr, gp(jb.wavenumber)[f], refph, col, part, TPC,
                                                           ngroups, KS, pe, smem, xs);
        __syncthreads();
        if constexpr (BF16X3 && PF)
            gemm_frequency_pf<CW, TT>(wave, lane, nK, NCG, prep_base + (int64_t)f * nUp * 8192, smem, pe, KS, BN, s_q);
        else
            gemm_frequency<PR, CW, TT>(wave, lane, nK, NCG, prep_base + (int64_t)f * nUp * 8192, smem, pe, KS, BN,
                                       s_q);
        __syncthreads();
        if constexpr (PR == kPrF16X2)
            store_frequency<kMfWaves>(jb, f, tile, BN, tid, NT, s_q, s_r0, s_help, 1, gp(jb.uscale)[f], s_xf);
        else
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
template <int PR, int VAR>
__global__ __launch_bounds__((kWsMf + kGenWaves) * 64) void bf_mfma_ws_kernel(const MfJob *__restrict__ jobs,
                                                                              int njobs)
{
    static_assert(PR != kPrF16X2, "the warp-specialised kernel has no f16x2 form");
    constexpr bool BF16X3 = PR == kPrBF16X3;
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
        s_help[part * BN + col] = steer_distances<kMfNS>(jb, valid, gx, gy, gz, part, TPC, ngroups, rr);
        if (part == 0) {
            double r0 = jb.ref_dist;
            if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
            s_r0[col] = r0;
        }
        for (int idx = gt; idx < 2 * kWsMf * BN; idx += NG) s_q[idx] = 0.f;
        const bool refph = jb.phase_mode == ABF_PHASE_REFERENCE;
        steer_generate<PR, (VAR & 1) != 0, (VAR >> 1), kMfNS>(rr, gp(jb.wavenumber)[phys(f0)], refph, col, part, TPC,
                                                           ngroups, KS, pe, xb(0));
        lds_barrier();
        for (int f = f0; f < f1; ++f) {
            const int b = (f - f0) & 1;
            if (f + 1 < f1)
                steer_generate<PR, (VAR & 1) != 0, (VAR >> 1), kMfNS>(rr, gp(jb.wavenumber)[phys(f + 1)], refph, col, part,
                                                                   TPC, ngroups, KS, pe, xb(b ^ 1));
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

} // namespace abf
