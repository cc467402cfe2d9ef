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

constexpr int kMfWaves = 8;
constexpr int kMfThreads = kMfWaves * 64;
constexpr int kMfNS = 5; // max 8-slot groups per thread in the steering generation

// One workgroup = 8 waves computes a tile of BN = 16 CT grid points against all Mp microphones
// for FB frequencies.  Prologue: float64 distances of the workgroup's (point, microphone) pairs
// into registers, steering-normalisation sums and r0.  Per frequency: the steering matrix X of
// the tile (BN x Mp, split hi+lo bf16 or fp32, K-slot order of kslot_mic) is generated into LDS;
// then each wave takes (row tile, column group) units in a boustrophedon order that balances
// the block triangle, streams its U fragments from L2 (prefetched one block ahead, each reused
// across CW column tiles) and accumulates Yr, Yi in MFMA registers; q = Re sum conj(x) Y is
// reduced per grid point in a fixed order (bit-identical results alone and batched), then
// normalised per steer type.
template <bool BF16X3, int CW>
__global__ __launch_bounds__(kMfThreads, 1) void bf_mfma_kernel(const MfJob *__restrict__ jobs, int njobs)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t wg = blockIdx.x;
    const MfJob &jb = jobs[find_job(jobs, njobs, wg)];

    const int M = jb.M, Mp = jb.Mp, CT = jb.CT, F = jb.F, FB = jb.FB;
    const int64_t G = jb.G;
    const int BN = 16 * CT;
    const int64_t local = wg - jb.wg_begin;
    const int tile = (int)(local % jb.ntiles);
    const int f0 = (int)(local / jb.ntiles) * FB;
    const int f1 = min(f0 + FB, F);
    const int KS = BF16X3 ? Mp + 8 : Mp + 4; // LDS row stride in elements (16-byte skew)
    const int nK = Mp / 32;
    const int nUp = nK * (nK + 1) / 2;

    // LDS: X planes [BN][KS] (bf16x3: re hi, re lo, im hi, im lo; fp32: re, im), then r0, help, q
    uint16_t *xh16 = reinterpret_cast<uint16_t *>(smem);
    float *x32 = reinterpret_cast<float *>(smem);
    const int64_t pe = (int64_t)BN * KS; // plane stride in elements
    const int64_t plane_bytes = BF16X3 ? pe * 2 : pe * 4;
    double *s_r0 = reinterpret_cast<double *>(smem + plane_bytes * (BF16X3 ? 4 : 2));
    double *s_help = s_r0 + BN;
    float *s_q = reinterpret_cast<float *>(s_help + BN); // [kMfWaves][BN]

    // ---- prologue: distances of this thread's (point, 8-slot groups) into registers ----
    const int TPC = kMfThreads / BN; // threads per grid point (power of two)
    const int col = tid / TPC, part = tid % TPC;
    const int64_t g = (int64_t)tile * BN + col;
    const bool valid = g < G;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        gx = jb.gpos[g];
        gy = jb.gpos[G + g];
        gz = jb.gpos[2 * G + g];
    }
    const int ngroups = Mp / 8; // group gi = 8 K-slots: block gi / 4, lane group gi % 4
    double rr[kMfNS][8];
    uint32_t live = 0; // bit s*8+e: slot holds a real microphone of a real grid point
    double help = 0.0;
#pragma unroll
    for (int s = 0; s < kMfNS; ++s) {
        const int gi = part + TPC * s;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int m = (gi >> 2) * 32 + kslot_mic(gi & 3, e);
            double r = 1.0;
            if (gi < ngroups && valid && m < M) {
                r = dist3(gx, gy, gz, jb.mpos[m], jb.mpos[M + m], jb.mpos[2 * M + m]);
                help += help_term(jb.steer_type, r);
                live |= 1u << (s * 8 + e);
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

    const int wk = weight_kind(jb.steer_type);
    const bool refph = jb.phase_mode == ABF_PHASE_REFERENCE;
    const bool fast = jb.fast_trig != 0;
    const int RT = Mp / 16;
    const int NCG = CT / CW;
    const int nunits = RT * NCG;
    const int rounds = (nunits + kMfWaves - 1) / kMfWaves;
    const char *__restrict__ prep_base = reinterpret_cast<const char *>(jb.prep);
    const u4 sg = u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};

    for (int f = f0; f < f1; ++f) {
        const double k = jb.wavenumber[f];
        // ---- steering matrix X of this frequency into LDS (wave-uniform selects only) ----
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
                    if (fast) {
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
            if constexpr (BF16X3) {
                const u4 *ap = reinterpret_cast<const u4 *>(ablk);
                u4 a_rh = ap[lane], a_rl = ap[64 + lane], a_ih = ap[128 + lane], a_il = ap[192 + lane];
                for (int J = I; J < nK; ++J) {
                    // prefetch the next block's fragments (blocks of a row are 8 KiB apart)
                    u4 n_rh = a_rh, n_rl = a_rl, n_ih = a_ih, n_il = a_il;
                    if (J + 1 < nK) {
                        const u4 *np = reinterpret_cast<const u4 *>(ablk + (int64_t)(J + 1 - I) * 8192);
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
                const f4 *ap = reinterpret_cast<const f4 *>(ablk);
                f4 u_a = ap[2 * lane], u_b = ap[2 * lane + 1], v_a = ap[128 + 2 * lane], v_b = ap[128 + 2 * lane + 1];
                for (int J = I; J < nK; ++J) {
                    f4 nu_a = u_a, nu_b = u_b, nv_a = v_a, nv_b = v_b;
                    if (J + 1 < nK) {
                        const f4 *np = reinterpret_cast<const f4 *>(ablk + (int64_t)(J + 1 - I) * 8192);
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
