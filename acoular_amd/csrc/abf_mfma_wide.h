// Wide streamed split-bf16 MFMA beamformer (device code), instantiated in abf_bf_mfma_wide.hip.
//
// The schedule of the streamed kernel (abf_mfma_stream.h): the K-blocks J of a frequency in
// descending order, the U blocks of each step staged through a two-slot LDS ring, the steering of
// the next block generated into the other of two LDS buffers while the MFMAs of the current block
// run, and the epilogue of row block J at the pair (J, J) that closes it.  What differs is the
// wave tile.  The streamed kernel runs 8 waves of one 16-column tile each, so every U fragment is
// read from LDS by the four waves of its row half and feeds 12 MFMAs per read; together with the
// steering planes that is about 2 MiB of LDS traffic per (frequency, 64-point tile), more than
// the CU's LDS delivers in the time of the tile's MFMAs.  Here 4 waves (one per SIMD) each own
// one 16-row half of every row block and two 16-column tiles (32 grid points):
//
//  * every U fragment read from LDS feeds 24 MFMAs, which halves the U reads;
//  * each thread generates 8 phasors per block (one 16-byte store per bf16 plane, two per
//    float32 plane) from float64 distances it keeps in registers for the whole band;
//  * with one wave per SIMD the accumulators (8 row blocks x 2 tiles x 8 values) and the
//    distances (8 blocks x 8) fit the 512 registers of a lane.
//
// Supported: split-bf16 form, 64 <= Mp <= 256 (NK = Mp / 32 blocks), hardware sin/cos with the
// reference's float32 phase; other cases use the alternating kernels.
#pragma once

#include <utility>

#include "abf_mfma_stream.h" // StSub, st_for, find_job, dist3

namespace abf {

template <int NK, int WK>
__global__ __launch_bounds__(kWdWaves * 64) __attribute__((amdgpu_waves_per_eu(1, 1))) void bf_mfma_wide_kernel(
    const MfJob *__restrict__ jobs, int njobs)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    uint16_t *xs = reinterpret_cast<uint16_t *>(smem);              // steering, bf16 [2][4][BN][KSB]
    float *xf = reinterpret_cast<float *>(smem + kWdOffXf);         // steering, f32 [2][2][BN][KSF]
    char *us = smem + kWdOffU;                                      // U ring [2][kWdGS][8 KiB]
    float *s_q = reinterpret_cast<float *>(smem + kWdOffQ);         // q partials [2][2][BN]
    double *s_fac = reinterpret_cast<double *>(smem + kWdOffFac);   // normalisation factor [BN]
    float *s_nst = reinterpret_cast<float *>(smem + kWdOffNst);     // steering normalisation [BN]
    double *s_r0 = reinterpret_cast<double *>(smem + kWdOffR0);     // [BN]
    double *s_help = reinterpret_cast<double *>(smem + kWdOffHelp); // steering-sum partials [4][BN]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t wg = blockIdx.x;
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];
    const int M = jb.M, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);

    // ---- generation role: grid point col = lane, K-slots 8 w .. 8 w + 7 of every block; slot
    // 8 g + e holds microphone 4 g + (e & 3) + 16 (e >> 2) (the K order of the prepared U) ----
    const int col = lane;
    const int64_t g = (int64_t)tile * kWdBN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = gp(jb.gpos)[g];
        gy = gp(jb.gpos)[G + g];
        gz = gp(jb.gpos)[2 * G + g];
    }
    const ABF_GLOBAL double *mp = gp(jb.mpos);
    double rr[NK][8];
    double help = 0.0;
#pragma unroll
    for (int J = 0; J < NK; ++J)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = 32 * J + 4 * wave + (e & 3) + 16 * (e >> 2);
            double r = 1.0;
            if (valid && m < M) {
                r = dist3(gx, gy, gz, mp[m], mp[M + m], mp[2 * M + m]);
                help += help_term(jb.steer_type, r);
            }
            rr[J][e] = r;
        }
    s_help[wave * kWdBN + col] = help;
    if (wave == 0) {
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0)) r0 = dist3(gx, gy, gz, jb.ref[0], jb.ref[1], jb.ref[2]);
        s_r0[col] = r0;
    }

    // phasors of a block (distances rb) for wavenumber k into steering buffer b: K-slots 8 w ..
    // 8 w + 7 of column col, as bf16 hi/lo planes (MFMA operand) and float32 (epilogue)
    auto generate = [&](const double (&rb)[8], double k, int b) __attribute__((always_inline)) {
        float xr[8], xi[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double r = rb[i];
            if constexpr (WK != 0) asm volatile("" : "+v"(r)); // keeps the weight in the band loop
            // the reference's float32 phase (fastFuncs.py:260), reduced to [-1/2, 1/2] revolution
            // in float64 for the hardware sin/cos
            const double phi = (double)(float)(k * r);
            const double u = phi * kInvTwoPi;
            const float t = (float)(u - rint(u));
            const float sn = __builtin_amdgcn_sinf(t);
            const float cs = __builtin_amdgcn_cosf(t);
            if constexpr (WK == 0) {
                xr[i] = cs;
                xi[i] = -sn;
            } else {
                const float rf = (float)r;
                const float w = WK == 1 ? rf : __builtin_amdgcn_rcpf(rf);
                xr[i] = w * cs;
                xi[i] = -w * sn;
            }
        }
        u4 rh, rl, ih, il;
        split8(xr, rh, rl);
        split8(xi, ih, il);
        uint16_t *dst = xs + b * kWdBlock + col * kWdKSB + 8 * wave;
        *reinterpret_cast<u4 *>(dst) = rh;
        *reinterpret_cast<u4 *>(dst + kWdPlane) = rl;
        *reinterpret_cast<u4 *>(dst + 2 * kWdPlane) = ih;
        *reinterpret_cast<u4 *>(dst + 3 * kWdPlane) = il;
        float *fd = xf + b * kWdBlockF + col * kWdKSF + 8 * wave;
        *reinterpret_cast<f4 *>(fd) = f4{xr[0], xr[1], xr[2], xr[3]};
        *reinterpret_cast<f4 *>(fd + 4) = f4{xr[4], xr[5], xr[6], xr[7]};
        *reinterpret_cast<f4 *>(fd + kWdPlaneF) = f4{xi[0], xi[1], xi[2], xi[3]};
        *reinterpret_cast<f4 *>(fd + kWdPlaneF + 4) = f4{xi[4], xi[5], xi[6], xi[7]};
    };

    // ---- MFMA role: row half p of every row block, column tiles 2 cp and 2 cp + 1 ----
    const int p = wave >> 1, cp = wave & 1;
    const int xoff = ((2 * cp) * 16 + (lane & 15)) * kWdKSB + (lane >> 4) * 8; // B fragment, tile 2 cp
    const int foff = ((2 * cp) * 16 + (lane & 15)) * kWdKSF + (lane >> 4) * 8 + p * 4; // epilogue values
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));
    constexpr int64_t kFreqBytes = (int64_t)(NK * (NK + 1) / 2) * 8192;
    constexpr StSub<NK, kWdGS> sub;
    constexpr int NS = sub.n;
    static_assert(NS % 2 == 0, "an even number of sub-steps per frequency");
    // this wave's 2 KiB share (bytes 2 KiB w .. 2 KiB w + 2 KiB) of U block (I, J) of frequency f
    auto load_share = [&](int f, int I, int J, int h) __attribute__((always_inline)) {
        const int ui = I * NK - (I * (I - 1)) / 2 + (J - I);
        return *reinterpret_cast<const ABF_GLOBAL u4 *>(prep_base + (int64_t)f * kFreqBytes + (int64_t)ui * 8192 +
                                                         wave * 2048 + h * 1024 + lane * 16);
    };
    // shares of the blocks of sub-step ts (of frequency ff) into a register set; the pairs of a
    // sub-step are I = I0 - q for q = 0 .. n - 1 (descending)
    auto load_sub = [&](auto tsc, int ff, u4 (&st)[kWdGS][2]) __attribute__((always_inline)) {
        constexpr int ts = decltype(tsc)::value;
#pragma unroll
        for (int q = 0; q < sub.I0[ts] - sub.I1[ts]; ++q) {
            st[q][0] = load_share(ff, sub.I0[ts] - q, sub.J[ts], 0);
            st[q][1] = load_share(ff, sub.I0[ts] - q, sub.J[ts], 1);
        }
    };
    auto write_sub = [&](auto tsc, int buf, const u4 (&st)[kWdGS][2]) __attribute__((always_inline)) {
        constexpr int ts = decltype(tsc)::value;
#pragma unroll
        for (int q = 0; q < sub.I0[ts] - sub.I1[ts]; ++q) {
            char *d = us + (buf * kWdGS + q) * 8192 + wave * 2048 + lane * 16;
            *reinterpret_cast<u4 *>(d) = st[q][0];
            *reinterpret_cast<u4 *>(d + 1024) = st[q][1];
        }
    };

    f4 accr[NK][2], acci[NK][2];
#pragma unroll
    for (int I = 0; I < NK; ++I)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            accr[I][c] = f4{0.f, 0.f, 0.f, 0.f};
            acci[I][c] = f4{0.f, 0.f, 0.f, 0.f};
        }
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    u4 stage[2][kWdGS][2];

    // first block of the band; U of its sub-step 0 into LDS buffer 0, of sub-step 1 into registers
    generate(rr[NK - 1], gp(jb.wavenumber)[f0], 0);
    load_sub(std::integral_constant<int, 0>{}, f0, stage[0]);
    load_sub(std::integral_constant<int, 1>{}, f0, stage[1]);
    write_sub(std::integral_constant<int, 0>{}, 0, stage[0]);
    __syncthreads();
    // frequency-independent normalisation of every grid point (fastFuncs.py:293-310):
    // result = q * fac, normalizeSteer = nst
    if (wave == 0) {
        double hs = 0.0;
        for (int q = 0; q < kWdWaves; ++q) hs += s_help[q * kWdBN + lane];
        double fac, nst;
        normalise(jb.steer_type, 1.0, s_r0[lane], hs, (double)M, jb.norm, fac, nst);
        s_fac[lane] = fac;
        s_nst[lane] = (float)nst;
    }
    __syncthreads();

    // maps of frequency ff from its q partials (the halves added in a fixed order)
    auto store = [&](int ff) __attribute__((always_inline)) {
        const float *sq = s_q + ((ff - f0) & 1) * 2 * kWdBN;
        const int64_t gg = (int64_t)tile * kWdBN + lane;
        if (gg < G) {
            const double qs = (double)sq[lane] + (double)sq[kWdBN + lane];
            float rf = (float)(qs * s_fac[lane]);
            if (jb.clip && rf < 0.f) rf = 0.f;
            gpw(jb.out)[(int64_t)ff * G + gg] = rf;
            if (jb.out_norm) gpw(jb.out_norm)[(int64_t)ff * G + gg] = s_nst[lane];
        }
    };

    int par = 0; // parity of the steering buffer in use at the first step of frequency f
    for (int f = f0; f < f1; ++f) {
        const int fn = min(f + 1, f1 - 1); // next frequency (clamped: the loads stay in bounds)
        const double knext = gp(jb.wavenumber)[fn];
        const double kcur = gp(jb.wavenumber)[f];
        // maps of the previous frequency: its q partials were written before the last barrier
        if (f > f0 && wave == 0) store(f - 1);
        float qacc[2] = {0.f, 0.f};
        u4 xrh[2], xrl[2], xih[2], xil[2], xnh[2], xnl[2];
        auto substep = [&](auto tc) __attribute__((always_inline)) {
            constexpr int t = decltype(tc)::value;
            constexpr int J = sub.J[t];
            constexpr int ub = t & 1; // LDS U buffer of this sub-step
            const int bcur = (par + (NK - 1 - J)) & 1;
            if constexpr (sub.first[t]) {
                // steering fragments of block J, column tiles 2 cp and 2 cp + 1
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const uint16_t *xb = xs + bcur * kWdBlock + xoff + c * 16 * kWdKSB;
                    xrh[c] = *reinterpret_cast<const u4 *>(xb);
                    xrl[c] = *reinterpret_cast<const u4 *>(xb + kWdPlane);
                    xih[c] = *reinterpret_cast<const u4 *>(xb + 2 * kWdPlane);
                    xil[c] = *reinterpret_cast<const u4 *>(xb + 3 * kWdPlane);
                    xnh[c] = xih[c] ^ sg;
                    xnl[c] = xil[c] ^ sg;
                }
            }
            // U of sub-step t + 1 (loaded one sub-step ago) into the LDS ring; loads of t + 2
            {
                constexpr int t1 = t + 1 < NS ? t + 1 : t + 1 - NS;
                constexpr int t2 = t + 2 < NS ? t + 2 : t + 2 - NS;
                write_sub(std::integral_constant<int, t1>{}, ub ^ 1, stage[(t + 1) & 1]);
                load_sub(std::integral_constant<int, t2>{}, t + 2 < NS ? f : fn, stage[t & 1]);
            }
            const char *ua = us + ub * (kWdGS * 8192) + p * 4096 + lane * 16;
#pragma unroll
            for (int q = 0; q < sub.I0[t] - sub.I1[t]; ++q) {
                constexpr int I0 = sub.I0[t];
                const int I = I0 - q;
                const char *a = ua + q * 8192;
                const u4 a_rh = *reinterpret_cast<const u4 *>(a);
                const u4 a_rl = *reinterpret_cast<const u4 *>(a + 1024);
                const u4 a_ih = *reinterpret_cast<const u4 *>(a + 2048);
                const u4 a_il = *reinterpret_cast<const u4 *>(a + 3072);
#define MF(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A), \
                                                                __builtin_bit_cast(bf16x8, B), C, 0, 0, 0)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    // Yr = Ur Xr - Ui Xi, Yi = Ui Xr + Ur Xi, each product hi*hi + hi*lo + lo*hi
                    MF(a_rh, xrh[c], accr[I][c]);
                    MF(a_ih, xrh[c], acci[I][c]);
                    MF(a_rh, xrl[c], accr[I][c]);
                    MF(a_ih, xrl[c], acci[I][c]);
                    MF(a_rl, xrh[c], accr[I][c]);
                    MF(a_il, xrh[c], acci[I][c]);
                    MF(a_ih, xnh[c], accr[I][c]);
                    MF(a_rh, xih[c], acci[I][c]);
                    MF(a_ih, xnl[c], accr[I][c]);
                    MF(a_rh, xil[c], acci[I][c]);
                    MF(a_il, xnh[c], accr[I][c]);
                    MF(a_rl, xih[c], acci[I][c]);
                }
#undef MF
                if (q == 0 && sub.first[t]) {
                    // pair (J, J) closes row block J: q += Re sum conj(x) Y over its rows 4 g + e of
                    // half p, whose float32 steering values sit at K-slots 8 g + 4 p + e
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const float *fb = xf + bcur * kWdBlockF + foff + c * 16 * kWdKSF;
                        const f4 xr = *reinterpret_cast<const f4 *>(fb);
                        const f4 xi = *reinterpret_cast<const f4 *>(fb + kWdPlaneF);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            qacc[c] = fmaf(xr[e], accr[J][c][e], qacc[c]);
                            qacc[c] = fmaf(xi[e], acci[J][c][e], qacc[c]);
                        }
                        accr[J][c] = f4{0.f, 0.f, 0.f, 0.f};
                        acci[J][c] = f4{0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
            if constexpr (sub.first[t]) {
                // next block: J - 1 of this frequency, or the last block of the next one
                if constexpr (J > 0)
                    generate(rr[J > 0 ? J - 1 : 0], kcur, bcur ^ 1);
                else
                    generate(rr[NK - 1], knext, bcur ^ 1);
            }
            if constexpr (J == 0 && sub.last[t]) {
                // the frequency's q partials of this wave's half and column tiles
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    float v = qacc[c];
                    v += __shfl_xor(v, 16);
                    v += __shfl_xor(v, 32);
                    if (lane < 16) s_q[((f - f0) & 1) * 2 * kWdBN + p * kWdBN + (2 * cp + c) * 16 + lane] = v;
                }
            }
            __syncthreads();
        };
        st_for(substep, std::make_integer_sequence<int, NS>{});
        par ^= NK & 1;
    }
    if (wave == 0) store(f1 - 1);
}

} // namespace abf
