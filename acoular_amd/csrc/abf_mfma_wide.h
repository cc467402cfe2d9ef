// Streamed split-bf16 MFMA beamformer for up to 512 microphones (device code), instantiated in
// abf_bf_mfma_wide.hip for NK = Mp / 32 in 2 .. 16.
//
// The schedule of the streamed kernel (abf_mfma_stream.h): the K-blocks J of a frequency in
// descending order; at step J every row block I <= J receives U(I, J) X(J), so row block J is
// complete at step J and its epilogue runs there; the U blocks of a step arrive by direct
// global-to-LDS loads into a three-slot ring two sub-steps ahead, each wave waiting before the
// barrier of a sub-step only for the loads of the next one; the steering of the next block is
// generated into the other of two LDS buffers after the MFMAs of the sub-step are issued.  The
// streamed kernel gives every wave one 16-row half of every row block, NK blocks x 8 accumulator
// registers plus NK x 8 registers of float64 distances: 256 registers at 512 microphones, more
// than a lane holds with two waves per SIMD.  Here the row blocks are split by parity instead:
//
//  * 8 waves over 32 grid points: wave w owns column tile c = w & 1 (16 points), row half
//    p = (w >> 1) & 1 and the row blocks of parity rgw = w >> 2, i.e. NK / 2 blocks x 8
//    accumulator registers; waves w and w + 4 (one SIMD) take the two parities of the same tile
//    and half, so every SIMD gets the MFMAs of every pair of a step for its tile and half;
//  * each thread generates, for the blocks of its parity, the 4 phasors per block its own
//    epilogue needs (column 16 c + (lane & 15), K-slots 8 (lane >> 4) + 4 p + i) from float64
//    distances it keeps for the band (NK / 2 x 4), and keeps their float32 values in registers
//    until the epilogue; a block is generated by the waves of its parity, so on every SIMD one
//    wave generates while the other issues MFMAs;
//  * each U fragment (half p of a pair) is read from LDS by the two column-tile waves of the
//    pair's parity and feeds 12 MFMAs per read, as in the streamed kernel.
//
// Supported: split-bf16 form, 64 <= Mp <= 512 (NK = Mp / 32 blocks), hardware sin/cos with the
// reference's float32 phase; other cases use the alternating kernels.
#pragma once

#include <utility>

#include "abf_mfma_stream.h" // StSub, st_for, find_job, dist3

namespace abf {

template <int NK, int WK>
__global__ __launch_bounds__(kWdWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void bf_mfma_wide_kernel(
    const MfJob *__restrict__ jobs, int njobs)
{
    constexpr int NH = (NK + 1) / 2; // row blocks of a parity (at most)
    extern __shared__ __attribute__((aligned(16))) char smem[];
    uint16_t *xs = reinterpret_cast<uint16_t *>(smem);              // steering, bf16 [2][4][BN][KSB]
    char *us = smem + kWdOffU;                                      // U ring [kWdRing][kWdGS][8 KiB]
    float *s_q = reinterpret_cast<float *>(smem + kWdOffQ);         // q partials [2][4][BN]
    double *s_fac = reinterpret_cast<double *>(smem + kWdOffFac);   // normalisation factor [BN]
    float *s_nst = reinterpret_cast<float *>(smem + kWdOffNst);     // steering normalisation [BN]
    double *s_r0 = reinterpret_cast<double *>(smem + kWdOffR0);     // [BN]
    double *s_help = reinterpret_cast<double *>(smem + kWdOffHelp); // steering-sum partials [4][BN]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = wave & 1, p = (wave >> 1) & 1, rgw = wave >> 2;
    const int grp = rgw * 2 + p; // q and steering-sum partial of this wave
    const int64_t wg = blockIdx.x;
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];
    const int M = jb.M, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);

    // ---- generation role: column col = 16 c + (lane & 15), K-slots 8 gq + 4 p + i of the blocks
    // J = 2 h + rgw (slot 8 g + e holds microphone 4 g + (e & 3) + 16 (e >> 2)) ----
    const int col = 16 * c + (lane & 15), gq = lane >> 4;
    const int64_t g = (int64_t)tile * kWdBN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = gp(jb.gpos)[g];
        gy = gp(jb.gpos)[G + g];
        gz = gp(jb.gpos)[2 * G + g];
    }
    const ABF_GLOBAL double *mp = gp(jb.mpos);
    double rr[NH][4];
    double help = 0.0;
#pragma unroll
    for (int h = 0; h < NH; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int J = 2 * h + rgw;
            const int m = 32 * J + 4 * gq + i + 16 * p;
            double r = 1.0;
            if (valid && J < NK && m < M) {
                r = dist3(gx, gy, gz, mp[m], mp[M + m], mp[2 * M + m]);
                help += help_term(jb.steer_type, r);
            }
            rr[h][i] = r;
        }
    // the four lane groups of a column, then the four wave groups (rgw, p) of a column in LDS
    help += __shfl_xor(help, 16);
    help += __shfl_xor(help, 32);
    if (gq == 0) s_help[grp * kWdBN + col] = help;
    if (wave == 0 && lane < kWdBN) {
        const int64_t g0 = (int64_t)tile * kWdBN + lane;
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0))
            r0 = g0 < G ? dist3(gp(jb.gpos)[g0], gp(jb.gpos)[G + g0], gp(jb.gpos)[2 * G + g0], jb.ref[0], jb.ref[1],
                                jb.ref[2])
                        : 1.0;
        s_r0[lane] = r0;
    }

    // float32 steering values of this thread's epilogue (re 0..3, im 4..7): one set, since the
    // waves of a parity close each of their blocks before they generate the next
    float ev[8];
    // phasors of distances rb for wavenumber k into steering buffer b
    auto generate = [&](const double (&rb)[4], double k, int b) __attribute__((always_inline)) {
        steer4<WK, true>(rb, k, ev, xs + b * kWdBlock + col * kWdKSB + 8 * gq + 4 * p, kWdPlane);
    };

    // ---- MFMA role: column tile c, row half p of the row blocks of parity rgw ----
    const int xoff = (c * 16 + (lane & 15)) * kWdKSB + (lane >> 4) * 8; // B fragment
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));
    constexpr int64_t kFreqBytes = (int64_t)(NK * (NK + 1) / 2) * 8192;
    constexpr StSub<NK, kWdGS, kWdRing> sub;
    constexpr int NS = sub.n;
    static_assert(NS % kWdRing == 0, "sub-steps per frequency a multiple of the ring slots");
    f4 accr[NH], acci[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        accr[h] = f4{0.f, 0.f, 0.f, 0.f};
        acci[h] = f4{0.f, 0.f, 0.f, 0.f};
    }
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    // this wave's 1 KiB share (half w >> 2, plane w & 3) of the U blocks of sub-step ts (pairs
    // I = I0 - q, descending) of frequency ff, straight from global memory into LDS buffer buf:
    // each lane loads 16 bytes, which land at the wave-uniform LDS address + 16 x lane
    auto dma_sub = [&](auto tsc, int ff, int buf) __attribute__((always_inline)) {
        constexpr int ts = decltype(tsc)::value;
#pragma unroll
        for (int q = 0; q < sub.I0[ts] - sub.I1[ts]; ++q) {
            const int I = sub.I0[ts] - q, J = sub.J[ts];
            const int ui = I * NK - (I * (I - 1)) / 2 + (J - I);
            const ABF_GLOBAL char *src =
                prep_base + (int64_t)ff * kFreqBytes + (int64_t)ui * 8192 + wave * 1024 + lane * 16;
            char *dst = us + (buf * kWdGS + q) * 8192 + wave * 1024;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                             (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
        }
    };

    const BandWavenumbers kof(jb.wavenumber, f0, f1, lane);
    // first block of the band (by the waves of its parity); U of its first kWdRing - 1 sub-steps
    // into LDS slots 0 ..
    if (rgw == ((NK - 1) & 1)) generate(rr[(NK - 1) >> 1], kof(f0), 0);
    dma_sub(std::integral_constant<int, 0>{}, f0, 0);
    if constexpr (kWdRing == 3) dma_sub(std::integral_constant<int, 1>{}, f0, 1);
    __syncthreads();
    // frequency-independent normalisation of every grid point (fastFuncs.py:293-310):
    // result = q * fac, normalizeSteer = nst
    if (wave == 0 && lane < kWdBN) {
        const double hs =
            (s_help[lane] + s_help[kWdBN + lane]) + (s_help[2 * kWdBN + lane] + s_help[3 * kWdBN + lane]);
        double fac, nst;
        normalise(jb.steer_type, 1.0, s_r0[lane], hs, (double)M, jb.norm, fac, nst);
        s_fac[lane] = fac;
        s_nst[lane] = (float)nst;
    }
    __syncthreads();

    // maps of frequency ff from its q partials (the four wave groups added in a fixed order)
    auto store = [&](int ff) __attribute__((always_inline)) {
        const float *sq = s_q + ((ff - f0) & 1) * 4 * kWdBN;
        const int64_t gg = (int64_t)tile * kWdBN + lane;
        if (lane < kWdBN && gg < G) {
            const double qs = ((double)sq[lane] + (double)sq[kWdBN + lane]) +
                              ((double)sq[2 * kWdBN + lane] + (double)sq[3 * kWdBN + lane]);
            float rf = (float)(qs * s_fac[lane]);
            if (jb.clip && rf < 0.f) rf = 0.f;
            gpw(jb.out)[(int64_t)ff * G + gg] = rf;
            if (jb.out_norm) gpw(jb.out_norm)[(int64_t)ff * G + gg] = s_nst[lane];
        }
    };

    int par = 0; // parity of the steering buffer in use at the first step of frequency f
    for (int f = f0; f < f1; ++f) {
        const int fn = min(f + 1, f1 - 1); // next frequency (clamped: the loads stay in bounds)
        const double kcur = kof(f), knext = kof(fn);
        // maps of the previous frequency: its q partials were written before the last barrier
        if (f > f0 && wave == 0) store(f - 1);
        float qacc = 0.f;
        u4 xrh, xrl, xih, xil, xnh, xnl;
        auto substep = [&](auto tc) __attribute__((always_inline)) {
            constexpr int t = decltype(tc)::value;
            constexpr int J = sub.J[t];
            constexpr int JG = J > 0 ? J - 1 : NK - 1; // block generated in this step
            constexpr int ub = t % kWdRing;                 // LDS U slot of this sub-step
            constexpr int tp = (t + kWdRing - 1) % NS;      // sub-step whose U loads are issued here
            constexpr int up = (t + kWdRing - 1) % kWdRing; // ... into this slot
            const int bcur = (par + (NK - 1 - J)) & 1;
            if constexpr (sub.first[t]) {
                // steering fragment of block J, column tile c (K-slots of lane group lane >> 4)
                if (J >= rgw) {
                    const uint16_t *xb = xs + bcur * kWdBlock + xoff;
                    xrh = *reinterpret_cast<const u4 *>(xb);
                    xrl = *reinterpret_cast<const u4 *>(xb + kWdPlane);
                    xih = *reinterpret_cast<const u4 *>(xb + 2 * kWdPlane);
                    xil = *reinterpret_cast<const u4 *>(xb + 3 * kWdPlane);
                    xnh = xih ^ sg;
                    xnl = xil ^ sg;
                }
            }
            // U of sub-step t + kWdRing - 1 (of the next frequency past the last) into its slot
            dma_sub(std::integral_constant<int, tp>{}, t + kWdRing - 1 < NS ? f : fn, up);
            // the MFMAs of this wave's pairs (parity rgw) of the sub-step, then the generation
            const char *ua = us + ub * (kWdGS * 8192) + p * 4096 + lane * 16;
#pragma unroll
            for (int q = 0; q < sub.I0[t] - sub.I1[t]; ++q) {
                const int I = sub.I0[t] - q;
                if ((I & 1) != rgw) continue;
                const char *a = ua + q * 8192;
                const u4 a_rh = *reinterpret_cast<const u4 *>(a);
                const u4 a_rl = *reinterpret_cast<const u4 *>(a + 1024);
                const u4 a_ih = *reinterpret_cast<const u4 *>(a + 2048);
                const u4 a_il = *reinterpret_cast<const u4 *>(a + 3072);
#define MF(A, B, C) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A), \
                                                                __builtin_bit_cast(bf16x8, B), C, 0, 0, 0)
                // Yr = Ur Xr - Ui Xi, Yi = Ui Xr + Ur Xi, each product hi*hi + hi*lo + lo*hi
                MF(a_rh, xrh, accr[I >> 1]);
                MF(a_ih, xrh, acci[I >> 1]);
                MF(a_rh, xrl, accr[I >> 1]);
                MF(a_ih, xrl, acci[I >> 1]);
                MF(a_rl, xrh, accr[I >> 1]);
                MF(a_il, xrh, acci[I >> 1]);
                MF(a_ih, xnh, accr[I >> 1]);
                MF(a_rh, xih, acci[I >> 1]);
                MF(a_ih, xnl, accr[I >> 1]);
                MF(a_rh, xil, acci[I >> 1]);
                MF(a_il, xnh, accr[I >> 1]);
                MF(a_rl, xih, acci[I >> 1]);
#undef MF
                if (q == 0 && sub.first[t]) {
                    // pair (J, J) closes row block J: q += Re sum conj(x) Y over its rows 4 g + i of
                    // half p, whose float32 steering values (K-slots 8 g + 4 p + i) this thread
                    // generated itself
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        qacc = fmaf(ev[e], accr[J >> 1][e], qacc);
                        qacc = fmaf(ev[4 + e], acci[J >> 1][e], qacc);
                    }
                    accr[J >> 1] = f4{0.f, 0.f, 0.f, 0.f};
                    acci[J >> 1] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
            if constexpr (sub.first[t]) {
                // the next block (J - 1 of this frequency, or the last block of the next one) by
                // the waves of its parity
                if (rgw == (JG & 1)) generate(rr[JG >> 1], J > 0 ? kcur : knext, bcur ^ 1);
            }
            if constexpr (J == 0 && sub.last[t]) {
                // the frequency's q partial of this wave's group and column tile
                qacc += __shfl_xor(qacc, 16);
                qacc += __shfl_xor(qacc, 32);
                if (lane < 16) s_q[((f - f0) & 1) * 4 * kWdBN + grp * kWdBN + c * 16 + lane] = qacc;
            }
            if constexpr (kWdRing == 2) {
                __syncthreads();
            } else {
                // the U loads of sub-step t + 1 (issued one sub-step earlier) have landed when at
                // most the ones issued in this sub-step are outstanding (a wave's loads complete in
                // order); the LDS writes and reads of this sub-step are done (lgkmcnt); then the
                // barrier.  The memory clobber keeps the compiler from moving LDS accesses across.
                constexpr int nd = sub.I0[tp] - sub.I1[tp];
                asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(nd) : "memory");
            }
        };
        st_for(substep, std::make_integer_sequence<int, NS>{});
        par ^= NK & 1;
    }
    // the U loads issued ahead for a next frequency land before the workgroup's LDS is released
    if constexpr (kWdRing == 3) __syncthreads();
    if (wave == 0) store(f1 - 1);
}

} // namespace abf
