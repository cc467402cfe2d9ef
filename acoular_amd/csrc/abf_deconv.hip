// Point spread function and DAMAS deconvolution for gfx950.
//
// PSF (calcPointSpreadFunction and its four formulations, acoular/fastFuncs.py:563-801):
//     psf[f][g][c] = |sum_m exp(-1j f32(k_f (r_gm - r_sm))) w_g(m) / r_sm|^2 * norm(g, s),  s = sources[c]
// with w_g(m) = 1 (classic), r_gm (inverse) or 1/r_gm (true level, true location).  The phase
// difference is rounded to float32 as the reference does before cos/sin (fastFuncs.py:671-673),
// so the sum does not factor into a product of per-point phasors and is evaluated term by term:
// one workgroup per 64 x 64 (grid point, source) tile and frequency, distances of both sides
// staged in LDS per chunk of microphones, 16 outputs per thread, float32 partial sums per chunk
// folded into float64 totals.
//
// DAMAS (damasSolverGaussSeidel, acoular/fastFuncs.py:820-855): one 1024-thread workgroup per
// frequency runs the relaxed Gauss-Seidel sweeps with the solution vector in LDS.  Rows go in
// blocks of 64: all 16 waves first form, for the 64 rows of a block, each row's sum over the
// columns that keep their current value while the block is updated (everything except the
// in-block columns up to the diagonal: columns before the block already hold this sweep's
// values, columns after it the last sweep's, exactly what the sequential sweep sees); one wave
// then walks the 64 rows in order, lane r adding A[r][l] x_l for every row l < r as soon as x_l
// is updated, and applies the relaxation and the positivity clip.  The update order is the
// reference's; only the float64 summation order differs.
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>

#include <cmath>
#include <limits>

#include "abf_common.h"
#include "abf_status.h"

namespace abf {

constexpr int kPsfT = 64;  // grid points and sources per tile
constexpr int kPsfMC = 32; // microphones per LDS chunk

struct PsfArgs {
    const double *gpos;       // [3][G]
    const double *mpos;       // [3][M]
    const double *wavenumber; // [F]
    const int64_t *sources;   // [S] grid indices
    void *out;                // [F][G][S] float or double
    double ref[3];
    double ref_dist;
    int64_t G, S;
    int M, steer_type, phase_mode, fast_trig, out_f64;
};

__device__ __forceinline__ double ref_distance(const PsfArgs &a, double x, double y, double z)
{
    return a.ref_dist > 0.0 ? a.ref_dist : dist3(x, y, z, a.ref[0], a.ref[1], a.ref[2]);
}

__global__ __launch_bounds__(256) void psf_kernel(PsfArgs a)
{
    __shared__ double s_rg[kPsfMC][kPsfT];
    __shared__ double s_rs[kPsfMC][kPsfT];
    __shared__ float s_wg[kPsfMC][kPsfT];
    __shared__ float s_ws[kPsfMC][kPsfT];
    __shared__ double s_help[4][kPsfT];
    __shared__ double s_r0g[kPsfT], s_r0s[kPsfT];
    __shared__ int s_srcok[kPsfT];

    const int tid = threadIdx.x;
    const int f = blockIdx.z;
    const int64_t g0 = (int64_t)blockIdx.x * kPsfT, c0 = (int64_t)blockIdx.y * kPsfT;
    const double k = a.wavenumber[f];
    const int M = a.M;

    // staging role: column col of both tiles, microphone rows mr, mr + 4, ...
    const int col = tid & (kPsfT - 1), mr = tid >> 6;
    const int64_t gst = g0 + col;
    const bool gok = gst < a.G;
    int64_t src = -1;
    if (c0 + col < a.S) {
        const int64_t s = a.sources[c0 + col];
        if (s >= 0 && s < a.G) src = s;
    }
    double gx = 0.0, gy = 0.0, gz = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    if (gok) {
        gx = a.gpos[gst];
        gy = a.gpos[a.G + gst];
        gz = a.gpos[2 * a.G + gst];
    }
    if (src >= 0) {
        sx = a.gpos[src];
        sy = a.gpos[a.G + src];
        sz = a.gpos[2 * a.G + src];
    }
    {
        double h = 0.0; // sum 1/r^2 over this thread's microphones (true level / location)
        for (int m = mr; m < M; m += 4) {
            const double r = dist3(gx, gy, gz, a.mpos[m], a.mpos[M + m], a.mpos[2 * M + m]);
            h += 1.0 / (r * r);
        }
        s_help[mr][col] = h;
        if (mr == 0) {
            s_r0g[col] = ref_distance(a, gx, gy, gz);
            s_r0s[col] = ref_distance(a, sx, sy, sz);
            s_srcok[col] = src >= 0;
        }
    }

    // output role: grid points tg + 16 i, sources ts + 16 j
    const int ts = tid & 15, tg = tid >> 4;
    double sr[4][4], si[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) sr[i][j] = si[i][j] = 0.0;

    for (int m0 = 0; m0 < M; m0 += kPsfMC) {
        const int mn = min(kPsfMC, M - m0);
        __syncthreads();
        for (int mm = mr; mm < mn; mm += 4) {
            const int m = m0 + mm;
            const double mx = a.mpos[m], my = a.mpos[M + m], mz = a.mpos[2 * M + m];
            const double rg = gok ? dist3(gx, gy, gz, mx, my, mz) : 1.0;
            const double rs = src >= 0 ? dist3(sx, sy, sz, mx, my, mz) : 1.0;
            float wg = 1.f;
            if (a.steer_type == ABF_STEER_INVERSE)
                wg = (float)rg;
            else if (a.steer_type != ABF_STEER_CLASSIC)
                wg = (float)(1.0 / rg);
            s_rg[mm][col] = rg;
            s_rs[mm][col] = rs;
            s_wg[mm][col] = wg;
            s_ws[mm][col] = (float)(1.0 / rs);
        }
        __syncthreads();
        float ar[4][4], ai[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) ar[i][j] = ai[i][j] = 0.f;
        for (int mm = 0; mm < mn; ++mm) {
            double rg[4], rs[4];
            float wg[4], ws[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                rg[i] = s_rg[mm][tg + 16 * i];
                wg[i] = s_wg[mm][tg + 16 * i];
                rs[i] = s_rs[mm][ts + 16 * i];
                ws[i] = s_ws[mm][ts + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float c, s;
                    phasor_rev(k * (rg[i] - rs[j]), a.phase_mode, a.fast_trig, c, s);
                    const float w = wg[i] * ws[j];
                    ar[i][j] = fmaf(w, c, ar[i][j]);
                    ai[i][j] = fmaf(-w, s, ai[i][j]);
                }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sr[i][j] += (double)ar[i][j];
                si[i][j] += (double)ai[i][j];
            }
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gl = tg + 16 * i;
        const int64_t g = g0 + gl;
        if (g >= a.G) continue;
        const double r0g = s_r0g[gl];
        const double help = s_help[0][gl] + s_help[1][gl] + s_help[2][gl] + s_help[3][gl];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int cl = ts + 16 * j;
            const int64_t c = c0 + cl;
            if (c >= a.S) continue;
            const double r0s = s_r0s[cl];
            const double sq = sr[i][j] * sr[i][j] + si[i][j] * si[i][j];
            double v;
            switch (a.steer_type) {
            case ABF_STEER_CLASSIC: {
                const double nf = r0s / M;
                v = sq * (nf * nf);
                break;
            }
            case ABF_STEER_INVERSE: {
                const double nf = r0s / r0g / M;
                v = sq * (nf * nf);
                break;
            }
            case ABF_STEER_TRUE_LEVEL: {
                const double nf = r0s / r0g / help;
                v = sq * (nf * nf);
                break;
            }
            default:
                v = sq * (r0s * r0s) / M / help;
                break;
            }
            if (!s_srcok[cl]) v = __builtin_nan("");
            const int64_t o = ((int64_t)f * a.G + g) * a.S + c;
            if (a.out_f64)
                static_cast<double *>(a.out)[o] = v;
            else
                static_cast<float *>(a.out)[o] = (float)v;
        }
    }
}

// -------------------------------------------------------------------------------------------
// DAMAS
// -------------------------------------------------------------------------------------------
constexpr int kDamasRB = 64;        // rows per block (one wave walks them in order)
constexpr int kDamasWaves = 16;     // 1024 threads
constexpr int64_t kDamasMaxG = 15360; // solution vector in LDS: 120 KiB of float64

struct DamasArgs {
    const void *A;     // [F][G][G] psf, float64 (a_f64) or float32
    const double *y;   // [F][G] dirty maps
    double *x;         // [F][G] start values in, solutions out
    int64_t G;
    int n_iter, a_f64;
    double relax;
};

// wave-wide sum through the device library's DPP reduction (no LDS round trips)
extern "C" __device__ double __ockl_wfred_add_f64(double);
__device__ __forceinline__ double wave_sum(double v) { return __ockl_wfred_add_f64(v); }

// lane l's value broadcast to the wave (l wave-uniform): two v_readlane, no LDS round trip
__device__ __forceinline__ double lane_value(double v, int l)
{
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l), hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <bool F64>
__device__ __forceinline__ double load_a(const void *A, int64_t idx)
{
    if constexpr (F64)
        return static_cast<const double *>(A)[idx];
    else
        return (double)static_cast<const float *>(A)[idx];
}

template <bool F64>
__global__ __launch_bounds__(1024) void damas_kernel(DamasArgs a)
{
    extern __shared__ double s_x[];                  // [G]
    __shared__ double s_blk[kDamasRB][kDamasRB + 1]; // in-block coefficients (row r, column l)
    __shared__ double s_part[kDamasRB];              // per row: sum over columns not yet updated
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t G = a.G;
    const int f = blockIdx.x;
    const double *y = a.y + (int64_t)f * G;
    double *x = a.x + (int64_t)f * G;
    const int64_t abase = (int64_t)f * G * G;
    const double w = a.relax;
    for (int64_t j = tid; j < G; j += blockDim.x) s_x[j] = x[j];
    __syncthreads();
    const int64_t nchunks = (G + 63) / 64;
    for (int it = 0; it < a.n_iter; ++it) {
        for (int64_t i0 = 0; i0 < G; i0 += kDamasRB) {
            const int nb = (int)min<int64_t>(kDamasRB, G - i0);
            // 1. for rows r = wave + 16 q: the sum over every column except the in-block ones at
            //    or left of the diagonal (those get this sweep's values in step 2), and the
            //    in-block coefficients; 4 column chunks of 64 per step keep 16 loads in flight.
            //    Loads are unconditional (row and column indices clamped into the matrix) and
            //    their values masked afterwards, so all of them are in flight before the first wait
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            {
                const int64_t j = i0 + lane, jc = min(j, G - 1);
                double bv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rr = min(wave + kDamasWaves * q, nb - 1);
                    bv[q] = load_a<F64>(a.A, abase + (i0 + rr) * G + jc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = wave + kDamasWaves * q;
                    if (r < nb) s_blk[r][lane] = j < G ? bv[q] : 0.0;
                }
            }
            for (int64_t t0 = 0; t0 < nchunks; t0 += 4) {
                double av[4][4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t j = (t0 + u) * 64 + lane, jc = min(j, G - 1);
                    xv[u] = s_x[jc];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int rr = min(wave + kDamasWaves * q, nb - 1);
                        av[q][u] = load_a<F64>(a.A, abase + (i0 + rr) * G + jc);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t j = (t0 + u) * 64 + lane;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r = wave + kDamasWaves * q;
                        const bool ok = j < G && r < nb && !(j >= i0 && j <= i0 + r);
                        acc[q] = fma(ok ? av[q][u] : 0.0, xv[u], acc[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = wave + kDamasWaves * q;
                const double sacc = wave_sum(acc[q]);
                if (r < nb && lane == 0) s_part[r] = sacc;
            }
            __syncthreads();
            // 2. one wave walks the rows in order; lane r keeps row r's running sum and adds
            //    A[r][l] x_l as soon as row l < r has its new value, so no step needs a reduction
            if (wave == 0) {
                const bool mine = lane < nb;
                double run = mine ? s_part[lane] : 0.0;
                const double xo = mine ? s_x[i0 + lane] : 0.0;
                const double yl = mine ? y[i0 + lane] : 0.0;
                double xn = xo;
                for (int l0 = 0; l0 < nb; l0 += 16) {
                    // the row's next 16 coefficients, read from LDS ahead of the dependent steps
                    double cf[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) cf[u] = s_blk[lane][l0 + u];
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        const int l = l0 + u;
                        if (l < nb) {
                            if (lane == l) {
                                const double v = (1.0 - w) * xo + w * (yl - run);
                                xn = v > 0.0 ? v : 0.0;
                            }
                            const double xl = lane_value(xn, l);
                            if (lane > l && mine) run = fma(cf[u], xl, run);
                        }
                    }
                }
                if (mine) s_x[i0 + lane] = xn;
            }
            __syncthreads();
        }
    }
    for (int64_t j = tid; j < G; j += blockDim.x) x[j] = s_x[j];
}

} // namespace abf

using namespace abf;

extern "C" {

int abf_psf(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, const int64_t *sources,
            int64_t num_sources, int32_t steer_type, int32_t phase_mode, int32_t out_f64, void *out, void *stream)
{
    if (!geom) return set_error(ABF_ERR_INVALID, "null geometry");
    if (steer_type < ABF_STEER_CLASSIC || steer_type > ABF_STEER_TRUE_LOCATION)
        return set_error(ABF_ERR_INVALID, "unknown steer_type");
    const int pm = phase_mode & 0xff;
    if (pm != ABF_PHASE_REFERENCE && pm != ABF_PHASE_FP64) return set_error(ABF_ERR_INVALID, "unknown phase_mode");
    const int64_t G = geom->num_grid;
    if (geom->num_mics < 1 || G < 0 || num_freqs < 0 || num_sources < 0)
        return set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_freqs > 65535) return set_error(ABF_ERR_INVALID, "at most 65535 frequencies per call");
    if ((num_sources + kPsfT - 1) / kPsfT > 65535) return set_error(ABF_ERR_INVALID, "too many sources per call");
    if (G == 0 || num_freqs == 0 || num_sources == 0) return ABF_OK;
    if (!geom->gpos || !geom->mpos || !wavenumber || !sources || !out) return set_error(ABF_ERR_INVALID, "null pointer");
    PsfArgs p{};
    p.gpos = geom->gpos;
    p.mpos = geom->mpos;
    p.wavenumber = wavenumber;
    p.sources = sources;
    p.out = out;
    for (int i = 0; i < 3; ++i) p.ref[i] = geom->ref[i];
    p.ref_dist = geom->ref_dist;
    p.G = G;
    p.S = num_sources;
    p.M = geom->num_mics;
    p.steer_type = steer_type;
    p.phase_mode = pm;
    p.fast_trig = (phase_mode & ABF_PHASE_FAST_TRIG) != 0;
    p.out_f64 = out_f64 != 0;
    const dim3 grid((unsigned)((G + kPsfT - 1) / kPsfT), (unsigned)((num_sources + kPsfT - 1) / kPsfT),
                    (unsigned)num_freqs);
    hipLaunchKernelGGL(psf_kernel, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

int abf_damas(const void *psf, int32_t psf_f64, const double *dirty, double *x, int64_t num_grid, int32_t num_freqs,
              int32_t n_iter, double relax, void *stream)
{
    if (num_grid < 0 || num_freqs < 0 || n_iter < 0) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_grid > kDamasMaxG) return set_error(ABF_ERR_UNSUPPORTED, "DAMAS solves grids of at most 15360 points");
    if (num_grid == 0 || num_freqs == 0 || n_iter == 0) return ABF_OK;
    if (!psf || !dirty || !x) return set_error(ABF_ERR_INVALID, "null pointer");
    DamasArgs d{psf, dirty, x, num_grid, n_iter, psf_f64 != 0, relax};
    const size_t lds = (size_t)num_grid * sizeof(double);
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)damas_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(kDamasMaxG * sizeof(double)));
        (void)hipFuncSetAttribute((const void *)damas_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(kDamasMaxG * sizeof(double)));
    });
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (psf_f64)
        hipLaunchKernelGGL(damas_kernel<true>, dim3((unsigned)num_freqs), dim3(1024), lds, s, d);
    else
        hipLaunchKernelGGL(damas_kernel<false>, dim3((unsigned)num_freqs), dim3(1024), lds, s, d);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
