// Streamed split-bf16 MFMA beamformer (device code), instantiated in abf_bf_mfma_stream.hip.
//
// The alternating kernels hold the whole steering matrix of a tile in LDS and run its generation
// (VALU: phase, reduction, cos/sin, hi/lo split) and the MFMA work one after the other, so the
// MFMA pipe idles during generation.  This kernel streams the steering matrix instead, one
// 32-microphone K-block at a time through a two-block LDS ring, and overlaps the generation of
// the next block with the MFMAs of the current one:
//
//  * the K-blocks J of a frequency are visited in descending order; at step J each wave adds
//    U(I, J) X(J) into the accumulators of every row block I <= J of its output tile.  Row block
//    J receives its last contribution at step J, so its epilogue q += Re sum conj(x) Y runs right
//    there, on the steering values of block J that the wave holds anyway (the K-slot order of
//    kslot_mic puts the values of result row r in the lane that holds row r);
//  * every wave owns one 16-column tile and one 16-row half of all row blocks (8 waves: 4
//    column tiles x 2 halves, 64 grid points per workgroup), so all waves carry the same MFMA
//    work in every step;
//  * each thread generates 4 phasors of the next block per step from float64 distances it keeps
//    in registers for the whole band (4 per block): the 4 values of its own column and row half
//    that its epilogue needs, whose float32 form it keeps in registers until then (16.27 ms on the
//    256-mic, 201 x 201, 513-bin workload with a float32 copy of the block in LDS instead, 15.66
//    ms so);
//  * the U blocks (both halves, 8 KiB) of the (I, J) pairs enter the workgroup once: a step is
//    cut into sub-steps of at most 4 pairs; at the start of sub-step t every wave issues its
//    1 KiB share of the blocks of sub-step t + 2 as direct global-to-LDS loads
//    (global_load_lds_dwordx4, no registers) into a three-slot LDS ring, and the barrier that ends
//    the sub-step hands over the slot of sub-step t + 1 after waiting for its loads only (15.52
//    against 15.76-15.84 ms with the shares staged through registers two sub-steps ahead; 14.29
//    against 14.62 ms with two slots, the loads of sub-step t + 1 issued in sub-step t).  The four
//    waves of a half read their fragments from LDS (global loads of the same fragment by four
//    waves would go through the CU's L1 four times, the bound of the first form of this kernel);
//  * the sign of -Ui Xi is taken from a negated copy of the steering fragment, made once per
//    step, instead of flipping the U fragment of every pair.
//
// Supported: split-bf16 form, 64 <= Mp <= 256 (NK = Mp / 32 blocks), hardware sin/cos with the
// reference's float32 phase; other cases use the alternating kernels.  The kernel runs either one
// work item (tile, frequency band) per workgroup or persistently (see bf_mfma_stream_kernel).
#pragma once

#include <utility>

#include "abf_mfma_impl.h" // store_frequency, find_job, dist3

namespace abf {

// fn(std::integral_constant<int, S>) for S = 0 .. N - 1, in order
template <class Fn, int... S> __device__ __forceinline__ void st_for(Fn &&fn, std::integer_sequence<int, S...>)
{
    (fn(std::integral_constant<int, S>{}), ...);
}

// The sub-steps of a frequency: step J (descending) covers the pairs I = J, J - 1, .., 0 (the
// pair (J, J) that closes row block J first), cut into ceil((J + 1) / GS) chunks of nearly
// equal size [I0, I1) counted downwards; the top step takes up to R - 1 more chunks so that the
// count is a multiple of R (which keeps the R slots of the LDS ring of U blocks in phase from one
// frequency to the next).
template <int NK, int GS = kStGS, int R = 2> struct StSub {
    static constexpr int kMax = NK * (NK + 1) / 2 + 2; // at least one pair per sub-step, one extra
    int n;
    int J[kMax], I0[kMax], I1[kMax];
    bool first[kMax], last[kMax];
    constexpr StSub() : n(0), J(), I0(), I1(), first(), last()
    {
        int tot = 0;
        for (int j = NK - 1; j >= 0; --j) tot += (j + GS) / GS;
        const int extra = (R - tot % R) % R;
        for (int j = NK - 1; j >= 0; --j) {
            const int np = j + 1;
            const int nc = (np + GS - 1) / GS + (j == NK - 1 ? extra : 0);
            int done = 0;
            for (int q = 0; q < nc; ++q) {
                const int sz = (np - done) / (nc - q) + ((np - done) % (nc - q) ? 1 : 0);
                J[n] = j;
                I0[n] = j - done;
                I1[n] = j - done - sz;
                first[n] = q == 0;
                last[n] = q == nc - 1;
                done += sz;
                ++n;
            }
        }
    }
};

// Four steering values of one grid point for wavenumber k from float64 distances rb (weight kind
// WK: 0 none, 1 r, 2 1/r): their float32 values into xe (re 0..3, im 4..7), their split-bf16 form
// (hi = RNE bf16, lo = RNE bf16 of the residual) into the four planes at dst (plane stride
// `plane` elements), 8 bytes per plane.  PIN keeps the weights from being hoisted out of the band
// loop (an opaque copy of the distance): the wide kernel needs it for its register budget; the
// streamed kernel runs faster without it (inverse steering 14.40-14.48 against 14.83-15.13 ms,
// true level 14.39-14.49 against 14.74-14.77 ms on the 256-mic, 201 x 201, 513-bin shape).
template <int WK, bool PIN>
__device__ __forceinline__ void steer4(const double (&rb)[4], double k, float (&xe)[8], uint16_t *dst, int plane)
{
    float *xr = xe, *xi = xe + 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double r = rb[i];
        if constexpr (WK != 0 && PIN) asm volatile("" : "+v"(r));
        // the reference's float32 phase (fastFuncs.py:260), reduced to [-1/2, 1/2] revolution in
        // float64 for the hardware sin/cos
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
    uint32_t h[4], l[4];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const float *src = d == 0 ? xr : xi;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bf16x2 hp = {(__bf16)src[2 * q], (__bf16)src[2 * q + 1]};
            uint32_t hb = __builtin_bit_cast(uint32_t, hp);
            // opaque to the compiler, which otherwise converts the first element a second time to
            // get its bf16 value back instead of shifting it out of the pair (13.99 against 14.08
            // ms on the 256-mic, 201 x 201, 513-bin workload)
            asm("" : "+v"(hb));
            const float h0 = __uint_as_float(hb << 16), h1 = __uint_as_float(hb & 0xffff0000u);
            const bf16x2 lp = {(__bf16)(src[2 * q] - h0), (__bf16)(src[2 * q + 1] - h1)};
            h[2 * d + q] = hb;
            l[2 * d + q] = __builtin_bit_cast(uint32_t, lp);
        }
    }
    *reinterpret_cast<uint2 *>(dst) = make_uint2(h[0], h[1]);
    *reinterpret_cast<uint2 *>(dst + plane) = make_uint2(l[0], l[1]);
    *reinterpret_cast<uint2 *>(dst + 2 * plane) = make_uint2(h[2], h[3]);
    *reinterpret_cast<uint2 *>(dst + 3 * plane) = make_uint2(l[2], l[3]);
}

// One work item: the tile and frequency band that workgroup wg of the one-item-per-workgroup
// launch computes.  Starts and ends with all LDS reads of the previous item done (barriers).
template <int NK, int WK>
__device__ __forceinline__ void stream_item(const MfJob *__restrict__ jobs, int njobs, int64_t wg, char *smem)
{
    uint16_t *xs = reinterpret_cast<uint16_t *>(smem);                   // steering, bf16 [2][4][BN][KSB]
    char *us = smem + kStOffU;                                           // U ring [kStRing][kStGS][8 KiB]
    float *s_q = reinterpret_cast<float *>(smem + kStOffQ);              // q partials [2][2][BN]
    double *s_fac = reinterpret_cast<double *>(smem + kStOffFac);        // normalisation factor [BN]
    float *s_nst = reinterpret_cast<float *>(smem + kStOffNst);          // steering normalisation [BN]
    double *s_r0 = reinterpret_cast<double *>(smem + kStOffR0);          // [BN]
    double *s_help = reinterpret_cast<double *>(smem + kStOffHelp);      // steering-sum partials [2][BN]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];
    const int M = jb.M, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);

    // ---- generation role: the values the thread's own epilogue needs: grid point col = 16 (wave & 3)
    // + (lane & 15), K-slots 8 gq + 4 pg + i (i < 4) of every block, gq = lane >> 4, pg = wave >> 2 ----
    const int col = 16 * (wave & 3) + (lane & 15), gq = lane >> 4, pg = wave >> 2;
    const int64_t g = (int64_t)tile * kStBN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = gp(jb.gpos)[g];
        gy = gp(jb.gpos)[G + g];
        gz = gp(jb.gpos)[2 * G + g];
    }
    const ABF_GLOBAL double *mp = gp(jb.mpos);
    double rr[NK][4];
    double help = 0.0;
#pragma unroll
    for (int J = 0; J < NK; ++J)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 32 * J + 4 * gq + i + 16 * pg;
            double r = 1.0;
            if (valid && m < M) {
                r = dist3(gx, gy, gz, mp[m], mp[M + m], mp[2 * M + m]);
                help += help_term(jb.steer_type, r);
            }
            rr[J][i] = r;
        }
    // the four lane groups of a column, then the two waves (pg) through LDS
    help += __shfl_xor(help, 16);
    help += __shfl_xor(help, 32);
    if (gq == 0) s_help[pg * kStBN + col] = help;
    if (wave == 0) {
        const int64_t g0 = (int64_t)tile * kStBN + lane;
        double r0 = jb.ref_dist;
        if (!(r0 > 0.0))
            r0 = g0 < G ? dist3(gp(jb.gpos)[g0], gp(jb.gpos)[G + g0], gp(jb.gpos)[2 * G + g0], jb.ref[0], jb.ref[1],
                                jb.ref[2])
                        : 1.0;
        s_r0[lane] = r0;
    }

    // float32 steering values of the epilogue, kept in registers from generation to epilogue: one
    // set, since the epilogue of block J runs before block J - 1 is generated (both in the first
    // sub-step of step J, MFMAs first)
    float ev[8];
    // phasors of a block (distances rb) for wavenumber k into steering buffer b: K-slots
    // 8 gq + 4 pg .. + 3 of column col, as bf16 hi/lo planes (MFMA operand) and float32 (epilogue)
    auto generate = [&](const double (&rb)[4], double k, int b, float (&xe)[8]) __attribute__((always_inline)) {
        // unpinned weights, except at 6 blocks, where the weighted instantiations would spill 7
        // registers without the pin
        steer4<WK, NK == 6>(rb, k, xe, xs + b * kStBlock + col * kStKSB + 8 * gq + 4 * pg, kStPlane);
    };

    // ---- MFMA role: column tile c, row half p of every row block ----
    const int c = wave & 3, p = wave >> 2;
    const int xoff = (c * 16 + (lane & 15)) * kStKSB + (lane >> 4) * 8;             // B fragment
    const ABF_GLOBAL char *prep_base = gp(reinterpret_cast<const char *>(jb.prep));
    constexpr int64_t kFreqBytes = (int64_t)(NK * (NK + 1) / 2) * 8192;
    constexpr StSub<NK, kStGS, kStRing> sub;
    constexpr int NS = sub.n;
    static_assert(NS % kStRing == 0, "sub-steps per frequency a multiple of the ring slots");
    f4 accr[NK], acci[NK];
#pragma unroll
    for (int I = 0; I < NK; ++I) {
        accr[I] = f4{0.f, 0.f, 0.f, 0.f};
        acci[I] = f4{0.f, 0.f, 0.f, 0.f};
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
            char *dst = us + (buf * kStGS + q) * 8192 + wave * 1024;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                             (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
        }
    };

    const BandWavenumbers kof(jb.wavenumber, f0, f1, lane);
    // first block of the band; U of its first kStRing - 1 sub-steps into LDS slots 0 ..
    generate(rr[NK - 1], kof(f0), 0, ev);
    dma_sub(std::integral_constant<int, 0>{}, f0, 0);
    if constexpr (kStRing == 3) dma_sub(std::integral_constant<int, 1>{}, f0, 1);
    __syncthreads();
    // frequency-independent normalisation of every grid point (fastFuncs.py:293-310):
    // result = q * fac, normalizeSteer = nst
    if (wave == 0) {
        double hs = 0.0;
        for (int q = 0; q < 2; ++q) hs += s_help[q * kStBN + lane];
        double fac, nst;
        normalise(jb.steer_type, 1.0, s_r0[lane], hs, (double)M, jb.norm, fac, nst);
        s_fac[lane] = fac;
        s_nst[lane] = (float)nst;
    }
    __syncthreads();

    // maps of frequency ff from its q partials (the halves added in a fixed order)
    auto store = [&](int ff) __attribute__((always_inline)) {
        const float *sq = s_q + ((ff - f0) & 1) * 2 * kStBN;
        const int64_t gg = (int64_t)tile * kStBN + lane;
        if (gg < G) {
            const double qs = (double)sq[lane] + (double)sq[kStBN + lane];
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
            constexpr int ub = t % kStRing;                 // LDS U slot of this sub-step
            constexpr int tp = (t + kStRing - 1) % NS;      // sub-step whose U loads are issued here
            constexpr int up = (t + kStRing - 1) % kStRing; // ... into this slot
            const int bcur = (par + (NK - 1 - J)) & 1;
            if constexpr (sub.first[t]) {
                // steering fragment of block J, column tile c (K-slots of lane group lane >> 4)
                const uint16_t *xb = xs + bcur * kStBlock + xoff;
                xrh = *reinterpret_cast<const u4 *>(xb);
                xrl = *reinterpret_cast<const u4 *>(xb + kStPlane);
                xih = *reinterpret_cast<const u4 *>(xb + 2 * kStPlane);
                xil = *reinterpret_cast<const u4 *>(xb + 3 * kStPlane);
                xnh = xih ^ sg;
                xnl = xil ^ sg;
            }
            // the steering generation of the next block, issued after this sub-step's MFMAs
            auto stage_gen = [&]() __attribute__((always_inline)) {
            if constexpr (sub.first[t]) {
                // next block: J - 1 of this frequency, or the last block of the next one
                if constexpr (J > 0)
                    generate(rr[J > 0 ? J - 1 : 0], kcur, bcur ^ 1, ev);
                else
                    generate(rr[NK - 1], knext, bcur ^ 1, ev);
            }
            };
            auto mfma = [&]() __attribute__((always_inline)) {
            const char *ua = us + ub * (kStGS * 8192) + p * 4096 + lane * 16;
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
                // Yr = Ur Xr - Ui Xi, Yi = Ui Xr + Ur Xi, each product hi*hi + hi*lo + lo*hi
                MF(a_rh, xrh, accr[I]);
                MF(a_ih, xrh, acci[I]);
                MF(a_rh, xrl, accr[I]);
                MF(a_ih, xrl, acci[I]);
                MF(a_rl, xrh, accr[I]);
                MF(a_il, xrh, acci[I]);
                MF(a_ih, xnh, accr[I]);
                MF(a_rh, xih, acci[I]);
                MF(a_ih, xnl, accr[I]);
                MF(a_rh, xil, acci[I]);
                MF(a_il, xnh, accr[I]);
                MF(a_rl, xih, acci[I]);
#undef MF
                if (q == 0 && sub.first[t]) {
                    // pair (J, J) closes row block J: q += Re sum conj(x) Y over its rows 4 g + i of
                    // half p, whose float32 steering values (K-slots 8 g + 4 p + i) this thread
                    // generated itself
                    const float *xe = ev;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        qacc = fmaf(xe[e], accr[J][e], qacc);
                        qacc = fmaf(xe[4 + e], acci[J][e], qacc);
                    }
                    accr[J] = f4{0.f, 0.f, 0.f, 0.f};
                    acci[J] = f4{0.f, 0.f, 0.f, 0.f};
                }
            }
            };
            // U of sub-step t + kStRing - 1 (of the next frequency past the last) into its slot
            dma_sub(std::integral_constant<int, tp>{}, t + kStRing - 1 < NS ? f : fn, up);
            // every wave issues its MFMAs first and generates while they drain: 14.14-14.66 ms on
            // the 256-mic, 201 x 201, 513-bin workload, against 15.22-15.89 ms with the two waves of
            // a SIMD in opposite orders and 15.24-15.30 ms with the generation first everywhere
            mfma();
            stage_gen();
            if constexpr (J == 0 && sub.last[t]) {
                // the frequency's q partial of this wave's half and column tile
                qacc += __shfl_xor(qacc, 16);
                qacc += __shfl_xor(qacc, 32);
                if (lane < 16) s_q[((f - f0) & 1) * 2 * kStBN + p * kStBN + c * 16 + lane] = qacc;
            }
            if constexpr (kStRing == 2) {
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
    if constexpr (kStRing == 3) __syncthreads();
    if (wave == 0) store(f1 - 1);
}

// PS = false: one work item per workgroup (workgroup index = item).  PS = true (persistent): a
// grid of at most one workgroup per CU walks the items 0 .. total - 1, workgroup b taking items
// b, b + grid, b + 2 grid, ... in order; every workgroup leaves once its next item passes total.
template <int NK, int WK, bool PS>
__global__ __launch_bounds__(kStWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void bf_mfma_stream_kernel(
    const MfJob *__restrict__ jobs, int njobs, int64_t total)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    if constexpr (!PS) {
        stream_item<NK, WK>(jobs, njobs, blockIdx.x, smem);
    } else {
        for (int64_t item = blockIdx.x; item < total; item += gridDim.x) {
            stream_item<NK, WK>(jobs, njobs, item, smem);
            __syncthreads();
        }
    }
}

} // namespace abf
