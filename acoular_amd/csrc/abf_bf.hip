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
//  * plain FP32 (bf_plain_kernel): one grid point per lane, U streamed through the scalar
//    cache as wave-uniform SGPR operands, the steering vectors of 64 grid points staged in LDS
//    per mic chunk;
//  * MFMA complex GEMM (bf_mfma_kernel): Y = U X on 16x16x32 bf16 MFMA with both operands
//    split hi+lo (3 products per real product) or on 16x16x4 fp32 MFMA; X (the steering
//    matrix of a grid tile) lives in LDS, U is pre-laid-out in MFMA fragment order by
//    prep_mfma_kernel and streamed from L2; q = Re sum_i conj(x_i) Y_i in the epilogue.
#include <type_traits>

#include "abf_mfma.h"

namespace abf {

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

// One grid point per lane, 64 per workgroup of 4 waves.  The microphones are taken in chunks of
// JC: the steering vectors of a chunk are generated into LDS ([JC][64] float2), then every row
// block of 16 rows i0 ... with i0 < chunk end accumulates (U v)_i over the chunk's columns
// j >= i; rows of earlier chunks take their v_i from a fresh evaluation.  The 16 values U_ij of
// row block i0 and column j (one 128-byte row of U^T) are wave-uniform: scalar loads into
// SGPRs, the operands of the FMAs, one column ahead of their use.  JC sets the LDS per workgroup
// (JC x 512 bytes) and so the workgroups per CU: JC = 64 takes 4 per CU at the cost of
// evaluating the steering vectors of earlier rows again (on 256 mics, 201 x 201, 513 bins:
// 99.6 ms for JC = 64, 105.8 ms for 128; the same loop with vector loads of U by every lane,
// which the L1 then serves 64 times over, 172 and 163 ms; scalar loads waited for at once,
// with JC = 256, 197 ms).  LU = true stages the U block of a row block and chunk in LDS instead
// (per wave [JC][16] float2, one 128-byte column per lane loaded with vector loads), every lane
// reading each value there (one broadcast address per read); the accumulation order is the same
// (115.6 ms against 99.3 ms on the workload above: each LDS read returns 16 bytes to all 64 lanes
// through the LDS data path, where one scalar load puts 64 bytes into SGPRs for the whole wave).
template <int JC, bool LU>
__global__ __launch_bounds__(256) void bf_plain_kernel(PlainArgs a)
{
    extern __shared__ float2 vlds[]; // [JC][kPlainNP], then (LU) [kPlainWaves][JC][kPlainRB]
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

    for (int c0 = 0; c0 < a.Mp; c0 += JC) {
        const int jn = min(JC, a.Mp - c0);
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
            const int js = max(i0, c0), je = c0 + jn;
            float ar[kPlainRB], ai[kPlainRB];
#pragma unroll
            for (int r = 0; r < kPlainRB; ++r) ar[r] = ai[r] = 0.f;
            // U through the scalar cache, one column ahead: the wait at the top of a column
            // completes column j's U row and v_j; column j + 1's loads are issued after it and
            // complete during column j's 64 FMAs (scalar loads return out of order, so any wait
            // for one of them is a wait for all: hence the explicit wait before the next loads)
            // (U^T is read-only in this kernel: read through the constant address space, the
            // loads stay scalar loads across the explicit waits)
            typedef const __attribute__((address_space(4))) float *cptr;
            auto ldu = [&](int j, float2 (&d)[kPlainRB]) __attribute__((always_inline)) {
                const cptr u = (cptr)(ut + (int64_t)j * a.Mp + i0);
#pragma unroll
                for (int r = 0; r < kPlainRB; ++r) d[r] = make_float2(u[2 * r], u[2 * r + 1]);
            };
            // ping-pong between two register sets, two columns per step: no copy between them
            // (a copy of a loaded value would wait for it at once)
            float2 ua[kPlainRB], ub[kPlainRB];
            float2 va = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
            auto fmas = [&](const float2 (&u)[kPlainRB], float2 v) __attribute__((always_inline)) {
#pragma unroll
                for (int r = 0; r < kPlainRB; ++r) {
                    ar[r] = fmaf(u[r].x, v.x, ar[r]);
                    ar[r] = fmaf(-u[r].y, v.y, ar[r]);
                    ai[r] = fmaf(u[r].x, v.y, ai[r]);
                    ai[r] = fmaf(u[r].y, v.x, ai[r]);
                }
            };
            if constexpr (LU) {
                float2 *uw = vlds + JC * kPlainNP + wave * (JC * kPlainRB);
                const int jl = js + lane;
                if (jl < je) {
                    const float4 *src = reinterpret_cast<const float4 *>(ut + (int64_t)jl * a.Mp + i0);
                    float4 *dst = reinterpret_cast<float4 *>(uw + (jl - js) * kPlainRB);
#pragma unroll
                    for (int t = 0; t < kPlainRB / 2; ++t) dst[t] = src[t];
                }
                // the LDS operations of one wave complete in order: these reads see the writes
                for (int j = js; j < je; ++j) {
                    const float4 *u4 = reinterpret_cast<const float4 *>(uw + (j - js) * kPlainRB);
                    float2 u[kPlainRB];
#pragma unroll
                    for (int t = 0; t < kPlainRB / 2; ++t) {
                        const float4 w = u4[t];
                        u[2 * t] = make_float2(w.x, w.y);
                        u[2 * t + 1] = make_float2(w.z, w.w);
                    }
                    fmas(u, vlds[(j - c0) * kPlainNP + lane]);
                }
            } else {
            if (js < je) {
                ldu(js, ua);
                va = vlds[(js - c0) * kPlainNP + lane];
            }
            for (int j = js; j < je; j += 2) {
                __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0): column j in ua, va
                __builtin_amdgcn_sched_barrier(0);
                const int j1 = min(j + 1, je - 1), j2 = min(j + 2, je - 1);
                ldu(j1, ub);
                vb = vlds[(j1 - c0) * kPlainNP + lane];
                fmas(ua, va);
                __builtin_amdgcn_sched_barrier(0);
                __builtin_amdgcn_s_waitcnt(0xc07f); // column j + 1 in ub, vb
                __builtin_amdgcn_sched_barrier(0);
                ldu(j2, ua);
                va = vlds[(j2 - c0) * kPlainNP + lane];
                fmas(ub, j + 1 < je ? vb : make_float2(0.f, 0.f)); // past the end: adds 0
                __builtin_amdgcn_sched_barrier(0);
            }
            }
            // a row block lies either inside the chunk (its v_i in LDS) or before it (i0 + 16 <= c0)
            if (i0 >= c0) {
#pragma unroll
                for (int r = 0; r < kPlainRB; ++r) {
                    const float2 vi = vlds[(i0 + r - c0) * kPlainNP + lane];
                    q = fmaf(vi.x, ar[r], q);
                    q = fmaf(vi.y, ai[r], q);
                }
            } else {
                // one row at a time (the accumulators shift down), so that one steering
                // evaluation is live at once
#pragma unroll 1
                for (int r = 0; r < kPlainRB; ++r) {
                    double h;
                    const float2 vi = plain_steer(a, p, i0 + r, k, h);
                    q = fmaf(vi.x, ar[0], q);
                    q = fmaf(vi.y, ai[0], q);
#pragma unroll
                    for (int t = 0; t + 1 < kPlainRB; ++t) {
                        ar[t] = ar[t + 1];
                        ai[t] = ai[t + 1];
                    }
                }
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

template __global__ void bf_plain_kernel<64, false>(PlainArgs);
template __global__ void bf_plain_kernel<128, false>(PlainArgs);
template __global__ void bf_plain_kernel<256, false>(PlainArgs);
template __global__ void bf_plain_kernel<64, true>(PlainArgs);

// -------------------------------------------------------------------------------------------
// float64 form: the plain kernel's structure with float64 steering vectors, CSM and sums, for
// callers that amplify input rounding (DAMAS): the reference's own arithmetic is float64 except
// for the float32 phase argument and the float32 cos/sin of it (fastFuncs.py:259-261), which is
// kept (cos/sin evaluated in float64 and rounded to float32, phase_mode REFERENCE)
// -------------------------------------------------------------------------------------------
__global__ void prep_plain64_kernel(const double2 *__restrict__ csm, int M, int Mp, int r_diag,
                                    double2 *__restrict__ ut)
{
    const int f = blockIdx.y;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)Mp * Mp) return;
    const int j = (int)(idx / Mp), i = (int)(idx % Mp);
    double2 v = make_double2(0.0, 0.0);
    if (i < M && j < M) {
        const double2 c = csm[((int64_t)f * M + i) * M + j];
        if (j > i)
            v = make_double2(2.0 * c.x, 2.0 * c.y);
        else if (j == i && !r_diag)
            v = make_double2(c.x, 0.0);
    }
    ut[((int64_t)f * Mp + j) * Mp + i] = v;
}

constexpr int kP64JC = 128; // microphones per LDS chunk ([128][64] double2 = 128 KiB)

__device__ __forceinline__ double2 plain64_steer(const Plain64Args &a, double gx, double gy, double gz, bool valid,
                                                 int m, double k, double &help)
{
    help = 0.0;
    if (!valid || m >= a.M) return make_double2(0.0, 0.0);
    const double r = dist3(gx, gy, gz, a.mpos[m], a.mpos[a.M + m], a.mpos[2 * a.M + m]);
    double phi = k * r, c, s;
    if (a.phase_mode == ABF_PHASE_REFERENCE) {
        phi = (double)(float)phi;
        sincos(phi, &s, &c);
        c = (double)(float)c;
        s = (double)(float)s;
    } else {
        sincos(phi, &s, &c);
    }
    double w = 1.0;
    if (a.steer_type == ABF_STEER_INVERSE) {
        w = r;
        help = r * r;
    } else if (a.steer_type != ABF_STEER_CLASSIC) {
        w = 1.0 / r;
        help = 1.0 / (r * r);
    }
    return make_double2(w * c, -w * s);
}

__global__ __launch_bounds__(256) void bf_plain64_kernel(Plain64Args a)
{
    extern __shared__ double2 vl64[]; // [kP64JC][kPlainNP]
    __shared__ double s_help[kPlainWaves][kPlainNP];
    __shared__ double s_q[kPlainWaves][kPlainNP];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int f = blockIdx.y;
    const int64_t g = (int64_t)blockIdx.x * kPlainNP + lane;
    const bool valid = g < a.G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = a.gpos[g];
        gy = a.gpos[a.G + g];
        gz = a.gpos[2 * a.G + g];
    }
    const double k = a.wavenumber[f];
    const double2 *__restrict__ ut = a.ut + (int64_t)f * a.Mp * a.Mp;
    double help = 0.0, q = 0.0;
    for (int c0 = 0; c0 < a.Mp; c0 += kP64JC) {
        const int jn = min(kP64JC, a.Mp - c0);
        __syncthreads();
        for (int mm = wave; mm < jn; mm += kPlainWaves) {
            double h;
            vl64[mm * kPlainNP + lane] = plain64_steer(a, gx, gy, gz, valid, c0 + mm, k, h);
            help += h;
        }
        __syncthreads();
        const int nrb = (c0 + jn) / kPlainRB;
        const int rounds = (nrb + kPlainWaves - 1) / kPlainWaves;
        for (int rd = 0; rd < rounds; ++rd) {
            const int rb = rd * kPlainWaves + ((rd & 1) ? (kPlainWaves - 1 - wave) : wave);
            if (rb >= nrb) continue;
            const int i0 = rb * kPlainRB;
            const int js = max(i0, c0);
            double ar[kPlainRB], ai[kPlainRB];
#pragma unroll
            for (int r = 0; r < kPlainRB; ++r) ar[r] = ai[r] = 0.0;
            for (int j = js; j < c0 + jn; ++j) {
                const double2 vj = vl64[(j - c0) * kPlainNP + lane];
                const double2 *__restrict__ u = ut + (int64_t)j * a.Mp + i0;
#pragma unroll
                for (int r = 0; r < kPlainRB; ++r) {
                    const double2 uu = u[r];
                    ar[r] = fma(uu.x, vj.x, ar[r]);
                    ar[r] = fma(-uu.y, vj.y, ar[r]);
                    ai[r] = fma(uu.x, vj.y, ai[r]);
                    ai[r] = fma(uu.y, vj.x, ai[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < kPlainRB; ++r) {
                const int i = i0 + r;
                double2 vi;
                if (i >= c0) {
                    vi = vl64[(i - c0) * kPlainNP + lane];
                } else {
                    double h;
                    vi = plain64_steer(a, gx, gy, gz, valid, i, k, h);
                }
                q = fma(vi.x, ar[r], q);
                q = fma(vi.y, ai[r], q);
            }
        }
    }
    s_q[wave][lane] = q;
    s_help[wave][lane] = help;
    __syncthreads();
    if (wave == 0 && valid) {
        double qs = 0.0, hs = 0.0;
#pragma unroll
        for (int w = 0; w < kPlainWaves; ++w) {
            qs += s_q[w][lane];
            hs += s_help[w][lane];
        }
        const double r0 = a.ref_dist > 0.0 ? a.ref_dist : dist3(gx, gy, gz, a.ref[0], a.ref[1], a.ref[2]);
        double res, nst;
        normalise(a.steer_type, qs, r0, hs, (double)a.M, a.norm, res, nst);
        if (a.clip && res < 0.0) res = 0.0;
        a.out[(int64_t)f * a.G + g] = res;
        if (a.out_norm) a.out_norm[(int64_t)f * a.G + g] = nst;
    }
}

// -------------------------------------------------------------------------------------------
// MFMA form
// -------------------------------------------------------------------------------------------
// Fragment layout of U for one frequency: upper 32x32 blocks (I <= J) in row-major upper
// order; per block two 16-row tiles; per tile 4 KiB:
//   bf16x3: parts {Ur_hi, Ur_lo, Ui_hi, Ui_lo}, each [64 lanes][8 bf16]
//   f16x2:  parts {Ur, -, Ui, -}, each [64 lanes][8 fp16], U scaled by 1 / uscale[f]
//   f32:    parts {Ur, Ui}, each [64 lanes][8 f32]
// Lane l of a tile holds row (l & 15) and columns 8*(l >> 4) + 0..7 of the block, the same
// k-set the B operand (X) uses, so each MFMA pairs matching k.

// f16x2 form: per frequency, the power of two s that maps the largest |Re U_ij|, |Im U_ij| of the
// upper triangle into [2^14, 2^15) (fp16 holds up to 65504); uscale[f] = 1 / s undoes it on q.
__global__ __launch_bounds__(256) void uscale_kernel(const float2 *__restrict__ csm, int M, int r_diag,
                                                     double *__restrict__ uscale)
{
    __shared__ float s_m[4];
    const int f = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    float mx = 0.f;
    const float2 *C = csm + (int64_t)f * M * M;
    // each wave walks rows i = wave, wave + 4, ...: the upper part j >= i of a row, lane by lane
    for (int i = tid >> 6; i < M; i += 4) {
        const float2 *row = C + (int64_t)i * M;
        for (int j = i + lane; j < M; j += 64) {
            if (j == i && r_diag) continue;
            const float2 c = row[j];
            const float a = j > i ? 2.f * fmaxf(fabsf(c.x), fabsf(c.y)) : fabsf(c.x);
            mx = fmaxf(mx, a); // fmaxf drops NaN: a NaN entry stays NaN in U and in the map
        }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((tid & 63) == 0) s_m[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        const float m = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
        int e = 0;
        if (m > 0.f && m < INFINITY) e = 14 - ilogbf(m); // m 2^e in [2^14, 2^15)
        uscale[f] = ldexp(1.0, -e);
    }
}

// One workgroup of 128 threads per (upper block (I, J), frequency): the 32 x 32 block of the CSM
// is read row-contiguously (32 threads per 256-byte row) into LDS, then thread (t, l) gathers the
// 8 entries of its fragment (row I 32 + 16 t + (l & 15), K-slots 8 (l >> 4) + e) from there.
__global__ __launch_bounds__(128) void prep_mfma_kernel(const float2 *__restrict__ csm, int M, int Mp, int r_diag,
                                                        int prec, const double *__restrict__ uscale,
                                                        char *__restrict__ prep)
{
    __shared__ float2 blk[32][33];
    const int nK = Mp / 32;
    // upper block u = upper_index(I, J, nK) of the grid's x index
    const int u = blockIdx.x;
    int I = 0;
    while (upper_index(I + 1, I + 1, nK) <= u) ++I;
    const int J = I + (u - upper_index(I, I, nK));
    const int f = blockIdx.y;
    const int k = threadIdx.x;
    const float2 *Cf = csm + (int64_t)f * M * M;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int r = q * 4 + (k >> 5), c = k & 31;
        const int i = I * 32 + r, j = J * 32 + c;
        blk[r][c] = (i < M && j < M) ? Cf[(int64_t)i * M + j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int t = k >> 6, l = k & 63;
    const int rl = t * 16 + (l & 15), i = I * 32 + rl;
    float ur[8], ui[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int cl = kslot_mic(l >> 4, e), j = J * 32 + cl;
        const float2 c = blk[rl][cl];
        float x = 0.f, y = 0.f;
        if (j > i) {
            x = 2.f * c.x;
            y = 2.f * c.y;
        } else if (j == i && !r_diag) {
            x = c.x;
        }
        ur[e] = x;
        ui[e] = y;
    }
    const int nUp = nK * (nK + 1) / 2;
    char *dst = prep + (((int64_t)f * nUp + upper_index(I, J, nK)) * 2 + t) * 4096;
    if (prec == kPrF16X2) {
        // U scaled by a power of two (exact) and rounded once to fp16: re into part 0, im into
        // part 2 of the fragment (parts 1 and 3 are not read by this form)
        const double sc = 1.0 / uscale[f];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            ur[e] = (float)((double)ur[e] * sc);
            ui[e] = (float)((double)ui[e] * sc);
        }
        u4 *d = reinterpret_cast<u4 *>(dst);
        d[0 * 64 + l] = pack8_h(ur);
        d[2 * 64 + l] = pack8_h(ui);
    } else if (prec == kPrBF16X3) {
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

// Eigenvalue form for every frequency of a job (the loop of BeamformerEig._calc,
// fbeamform.py:852-888, around _freqBeamformer_EigValues, fastFuncs.py:335-411):
//     q = sum_{e in [e0, e0+K)} l_e (|v_e^H h|^2 - [r_diag] sum_m |v_me|^2 |h_m|^2)
// One workgroup per (P grid points, frequency).  The steering vectors of the P points for all
// M microphones are generated once into LDS ([M][P] complex float32); thread t then takes grid
// point t % P and eigenvectors e0 + t / P, e0 + t / P + 256 / P, ..., so every steering element
// is reused by all K eigenvectors.  P (a power of two, 4..64) is chosen on the host so that the
// LDS tile fits.
__global__ __launch_bounds__(256) void bf_eig_grid_kernel(EigGridArgs a)
{
    extern __shared__ float2 hl[]; // [M][P]
    __shared__ double s_red[256];
    const int tid = threadIdx.x;
    const int P = a.P;
    const int p = tid & (P - 1);
    const int slot = tid / P;
    const int nslot = 256 / P;
    const int f = blockIdx.y;
    const int64_t g = (int64_t)blockIdx.x * P + p;
    const bool valid = g < a.G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = a.gpos[g];
        gy = a.gpos[a.G + g];
        gz = a.gpos[2 * a.G + g];
    }
    const double k = a.wavenumber[f];
    double help = 0.0;
    for (int m = slot; m < a.M; m += nslot) {
        float2 v = make_float2(0.f, 0.f);
        if (valid) {
            const double r = dist3(gx, gy, gz, a.mpos[m], a.mpos[a.M + m], a.mpos[2 * a.M + m]);
            float c, s, w;
            double h;
            phasor_rev(k * r, a.phase_mode, a.fast_trig, c, s);
            steer_weight(a.steer_type, r, w, h);
            help += h;
            v = make_float2(w * c, -w * s);
        }
        hl[m * P + p] = v;
    }
    s_red[tid] = help;
    __syncthreads();
    double hs = 0.0;
    if (tid < P)
        for (int i = 0; i < nslot; ++i) hs += s_red[tid + P * i];
    __syncthreads();

    const float2 *__restrict__ V = a.eigvec + (int64_t)f * a.M * a.M;
    const double *__restrict__ lam = a.eigval + (int64_t)f * a.M;
    double q = 0.0;
    for (int e = a.e0 + slot; e < a.e0 + a.K; e += nslot) {
        float tr = 0.f, ti = 0.f, ds = 0.f;
        for (int m = 0; m < a.M; ++m) {
            const float2 ev = V[(int64_t)m * a.M + e]; // conj(V_me) * h_m
            const float2 h = hl[m * P + p];
            const float pr = ev.x * h.x + ev.y * h.y;
            const float pi = ev.x * h.y - ev.y * h.x;
            tr += pr;
            ti += pi;
            ds = fmaf(pr, pr, fmaf(pi, pi, ds));
        }
        double full = (double)tr * tr + (double)ti * ti;
        if (a.r_diag) full -= ds;
        q += full * lam[e];
    }
    s_red[tid] = q;
    __syncthreads();
    if (tid < P && valid) {
        double qs = 0.0;
        for (int i = 0; i < nslot; ++i) qs += s_red[tid + P * i];
        const double r0 = a.ref_dist > 0.0 ? a.ref_dist : dist3(gx, gy, gz, a.ref[0], a.ref[1], a.ref[2]);
        double res, nst;
        normalise(a.steer_type, qs, r0, hs, (double)a.M, a.norm, res, nst);
        float rf = (float)res;
        if (a.clip && rf < 0.f) rf = 0.f;
        a.out[(int64_t)f * a.G + g] = rf;
        if (a.out_norm) a.out_norm[(int64_t)f * a.G + g] = (float)nst;
    }
}

// The same eigenvalue form on fp32 MFMA when many eigenvectors are summed (BeamformerMusic,
// BeamformerFunctional): Y = V_K^H H is a complex GEMM (K x M times M x P) and
// q = sum_k l_k |Y_kg|^2.  One workgroup per (P grid points, frequency) generates H into LDS as
// bf_eig_grid_kernel does; each wave then takes (16-eigenvector, 16-point) tiles, streams the
// eigenvector fragments from L2 (lane l: eigenvector l % 16, microphone l / 16 of a 4-microphone
// step, 64 contiguous bytes per 16 lanes) and the steering fragments from LDS, accumulates
// Re Y and Im Y with four v_mfma_f32_16x16x4_f32 per step, and folds l_k |Y|^2 of its 16 rows
// into per-point sums.  float32 products and sums as in the VALU kernel; without r_diag only
// (the r_diag correction needs per-microphone weights of the whole range; that case stays on
// the VALU kernel).
__global__ __launch_bounds__(256) void bf_eig_mfma_kernel(EigGridArgs a)
{
    extern __shared__ float2 hx[]; // [Mp][P], microphones padded to a multiple of 4
    __shared__ double s_red[256];
    __shared__ double s_part[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = a.P;
    const int Mp = (a.M + 3) & ~3;
    const int f = blockIdx.y;
    const int p = tid & (P - 1);
    const int slot = tid / P, nslot = 256 / P;
    const int64_t g = (int64_t)blockIdx.x * P + p;
    const bool valid = g < a.G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = a.gpos[g];
        gy = a.gpos[a.G + g];
        gz = a.gpos[2 * a.G + g];
    }
    const double k = a.wavenumber[f];
    double help = 0.0;
    for (int m = slot; m < Mp; m += nslot) {
        float2 v = make_float2(0.f, 0.f);
        if (valid && m < a.M) {
            const double r = dist3(gx, gy, gz, a.mpos[m], a.mpos[a.M + m], a.mpos[2 * a.M + m]);
            float c, s, w;
            double h;
            phasor_rev(k * r, a.phase_mode, a.fast_trig, c, s);
            steer_weight(a.steer_type, r, w, h);
            help += h;
            v = make_float2(w * c, -w * s);
        }
        hx[m * P + p] = v;
    }
    s_red[tid] = help;
    if (tid < 4 * 64) s_part[tid >> 6][tid & 63] = 0.0;
    __syncthreads();
    double hs = 0.0;
    if (tid < P)
        for (int i = 0; i < nslot; ++i) hs += s_red[tid + P * i];

    const float2 *__restrict__ V = a.eigvec + (int64_t)f * a.M * a.M;
    const double *__restrict__ lam = a.eigval + (int64_t)f * a.M;
    const int nrt = (a.K + 15) / 16, nct = P / 16;
    const int row = lane & 15, grp = lane >> 4;
    for (int t = wave; t < nrt * nct; t += 4) {
        const int rt = t / nct, ct = t % nct;
        const int e = a.e0 + rt * 16 + row; // this lane's eigenvector in the A fragment
        const bool eok = e < a.e0 + a.K;
        const int ec = eok ? e : a.e0;
        // two accumulator pairs (even and odd 4-microphone steps) halve the dependent MFMA chains
        f4 accr[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
        f4 acci[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
        // eigenvector values of microphones m0 + 4 u + grp, u < 4: loaded one pass ahead,
        // unconditionally from a clamped row, and zeroed outside the array only after the
        // current pass's MFMAs, so that the loads stay in flight across them
        auto load_ev = [&](int m0, float2 (&ev)[4]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < 4; ++u) ev[u] = V[(int64_t)min(m0 + 4 * u + grp, a.M - 1) * a.M + ec];
        };
        auto mask_ev = [&](int m0, float2 (&ev)[4]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (!(m0 + 4 * u + grp < a.M && eok)) ev[u] = make_float2(0.f, 0.f);
        };
        float2 ev[4];
        load_ev(0, ev);
        mask_ev(0, ev);
        for (int m0 = 0; m0 < Mp; m0 += 16) {
            float2 nev[4], hv[4];
            load_ev(m0 + 16, nev);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int m = m0 + 4 * u + grp;
                hv[u] = m < Mp ? hx[m * P + ct * 16 + row] : make_float2(0.f, 0.f);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                // Y = conj(V)^T H: Re = Vr Hr + Vi Hi, Im = Vr Hi - Vi Hr
                f4 &ar = accr[u & 1], &ai = acci[u & 1];
                ar = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[u].x, hv[u].x, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[u].y, hv[u].y, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[u].x, hv[u].y, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_16x16x4f32(-ev[u].y, hv[u].x, ai, 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
            mask_ev(m0 + 16, nev);
#pragma unroll
            for (int u = 0; u < 4; ++u) ev[u] = nev[u];
        }
        const f4 yr = accr[0] + accr[1], yi = acci[0] + acci[1];
        // lane (col = l % 16, group) holds rows 4 group + i of the tile
        double sq = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int er = a.e0 + rt * 16 + 4 * grp + i;
            if (er < a.e0 + a.K) sq += ((double)yr[i] * yr[i] + (double)yi[i] * yi[i]) * lam[er];
        }
        sq += __shfl_xor(sq, 16, 64);
        sq += __shfl_xor(sq, 32, 64);
        if (grp == 0) s_part[wave][ct * 16 + row] += sq;
    }
    __syncthreads();
    if (tid < P && valid) {
        const double qs = s_part[0][tid] + s_part[1][tid] + s_part[2][tid] + s_part[3][tid];
        const double r0 = a.ref_dist > 0.0 ? a.ref_dist : dist3(gx, gy, gz, a.ref[0], a.ref[1], a.ref[2]);
        double res, nst;
        normalise(a.steer_type, qs, r0, hs, (double)a.M, a.norm, res, nst);
        float rf = (float)res;
        if (a.clip && rf < 0.f) rf = 0.f;
        a.out[(int64_t)f * a.G + g] = rf;
        if (a.out_norm) a.out_norm[(int64_t)f * a.G + g] = (float)nst;
    }
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
