// CLEAN-SC (BeamformerCleansc._calc, acoular/fbeamform.py:1449-1492) and CLEAN
// (BeamformerClean._calc, fbeamform.py:1536-1598) for gfx950.
//
// CLEAN-SC: the reference runs, per frequency, up to J iterations of: peak of the map h, coherent
// source component hh of the CSM at the peak (a 20-step fixed point on the steering vector),
// removal of hh's map from h and of hmax hh hh^H from the CSM.  Here every frequency of a request
// advances together, two launches per iteration:
//  * csc_select_kernel, one workgroup per frequency: the peak (first index of the maximum, as
//    numpy's argmax), the float32 result update, the stop test, the steering vector at the peak
//    (transfer function, SteeringVector.steer_vector normalisation, fbeamform.py:145-153,
//    207-258), D1 = (C^T - diag C) w / hmax, the fixed point, and the CSM update;
//  * csc_map_kernel, one thread per (grid point, frequency): h -= damp * (the eigenvalue-form map
//    of the single pair (hmax, conj(hh)), fastFuncs.py:335-411).
// A frequency whose stop test fires is skipped by all later launches, so the result equals the
// reference's early break.  Arithmetic is float64 except the reference's float32 phases and
// float32 cos/sin (fastFuncs.py:259-261, 896-897).
//
// CLEAN (Hoegbom): per frequency, the peak of the dirty map, its amplitude damp * max added to
// the clean map, and the PSF column of the peak (calcPointSpreadFunction, fastFuncs.py:563-801,
// float32 phase differences) times that amplitude subtracted from the dirty map, until an
// iteration no longer lowers the sum of absolute values, n_iter iterations have run or no
// positive value is left.  Again two launches per iteration for all frequencies:
// clean_select_kernel (sum, peak, stop test, amplitude; one workgroup per frequency) and
// clean_update_kernel (the PSF column term by term and the subtraction; one thread per grid
// point and frequency).
//
// Both methods read the distances grid point -> microphone and the steering weights from a
// per-call table ([M][G], csc_table_kernel) instead of recomputing them in every iteration.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "abf_common.h"
#include "abf_status.h"

namespace abf {

constexpr int kCscMaxM = 1024;
constexpr int kCscThreads = 256;
constexpr int kCscPer = kCscMaxM / kCscThreads; // microphones per thread in the selection kernel
constexpr int kCscBatch = 8;                     // loads issued together before their first use

struct CscArgs {
    const double *gpos;       // [3][G]
    const double *mpos;       // [3][M]
    const double *wavenumber; // [F]
    double2 *csm;             // [F][M][M] working copy, updated
    double *h;                // [F][G] map, updated
    float *result;            // [F][G]
    double *powers;           // [F][J]
    double2 *hh;              // [F][M]
    double *hmax;             // [F]
    int *stopped;             // [F]
    double2 *rw;              // [M][G]: distance grid point -> mic, steering weight
    double *help;             // [G]: the steer type's sum over the mics (fastFuncs.py:361-372)
    double ref[3];
    double ref_dist;
    double norm, damp;
    int64_t G;
    int M, F, J, stopn;
    int steer_type, r_diag, phase_mode;
};

__device__ __forceinline__ double csc_r0(const CscArgs &a, double gx, double gy, double gz)
{
    return a.ref_dist > 0.0 ? a.ref_dist : dist3(gx, gy, gz, a.ref[0], a.ref[1], a.ref[2]);
}

extern "C" __device__ double __ockl_wfred_add_f64(double);

// Sum over the 256 threads of a workgroup with one barrier: wave sums, then the four partials in
// wave order.  red holds two sets of four slots used alternately (par flips on every call), so a
// call never overwrites slots that a thread may still be reading from the previous call.
__device__ __forceinline__ double block_sum(double v, double *red, int &par)
{
    v = __ockl_wfred_add_f64(v);
    double *r = red + 4 * par;
    par ^= 1;
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((r[0] + r[1]) + r[2]) + r[3];
}

__global__ __launch_bounds__(256) void fill_zero_f32_kernel(float *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ __launch_bounds__(256) void fill_zero_i32_kernel(int *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Per call: distances and steering weights of every (mic, grid point) and the per-point sums of
// the steer normalisation (sum r^2 for inverse, sum 1/r^2 for true level and true location), in
// the microphone order of the reference's loops.
__global__ __launch_bounds__(256) void csc_table_kernel(const double *gpos, const double *mpos, double2 *rw,
                                                        double *help_out, int64_t G, int M, int steer_type)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const double gx = gpos[g], gy = gpos[G + g], gz = gpos[2 * G + g];
    double help = 0.0;
    for (int m = 0; m < M; ++m) {
        const double r = dist3(gx, gy, gz, mpos[m], mpos[M + m], mpos[2 * M + m]);
        double w = 1.0;
        if (steer_type == ABF_STEER_INVERSE) {
            w = r;
            help += r * r;
        } else if (steer_type != ABF_STEER_CLASSIC) {
            w = 1.0 / r;
            help += 1.0 / (r * r);
        }
        rw[(int64_t)m * G + g] = make_double2(r, w);
    }
    help_out[g] = help;
}

__global__ __launch_bounds__(kCscThreads) void csc_select_kernel(CscArgs a, int j)
{
    __shared__ double s_pv[4];
    __shared__ int64_t s_pi[4];
    __shared__ double s_red[8];
    __shared__ double2 s_w[kCscMaxM], s_hh[kCscMaxM], s_part[kCscThreads];
    __shared__ int s_stop;
    const int f = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int M = a.M;
    const int64_t G = a.G;
    if (a.stopped[f]) return;
    int par = 0;
    // 1. the peak of h: largest value, first index on ties (numpy's argmax)
    const double *h = a.h + (int64_t)f * G;
    double best = -INFINITY;
    int64_t bi = G;
    for (int64_t base = tid; base < G; base += kCscBatch * kCscThreads) {
        double v[kCscBatch];
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t g = base + (int64_t)u * kCscThreads;
            v[u] = h[g < G ? g : G - 1];
        }
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t g = base + (int64_t)u * kCscThreads;
            if (g < G && v[u] > best) {
                best = v[u];
                bi = g;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double v = __shfl_xor(best, off);
        const int64_t i = __shfl_xor(bi, off);
        if (v > best || (v == best && i < bi)) {
            best = v;
            bi = i;
        }
    }
    if (lane == 0) {
        s_pv[wv] = best;
        s_pi[wv] = bi;
    }
    __syncthreads();
    double hmax = s_pv[0];
    int64_t xi = s_pi[0];
    for (int q = 1; q < 4; ++q) {
        const double v = s_pv[q];
        const int64_t i = s_pi[q];
        if (v > hmax || (v == hmax && i < xi)) {
            hmax = v;
            xi = i;
        }
    }
    if (xi >= G) xi = 0;
    // 2. bookkeeping and the stop test (the break comes after the result update)
    if (tid == 0) {
        a.powers[(int64_t)f * a.J + j] = hmax;
        float *res = a.result + (int64_t)f * G + xi;
        *res = (float)((double)*res + a.damp * hmax);
        s_stop = (j > a.stopn && hmax > a.powers[(int64_t)f * a.J + j - a.stopn]) ? 1 : 0;
        if (s_stop) a.stopped[f] = 1;
    }
    __syncthreads();
    if (s_stop) return;
    // 3. steering vector at the peak: transfer a_m = exp(-1j f32(k (r_m - r0))) r0 / r_m
    const double k = a.wavenumber[f];
    const double gx = a.gpos[xi], gy = a.gpos[G + xi], gz = a.gpos[2 * G + xi];
    const double r0 = csc_r0(a, gx, gy, gz);
    double2 wr[kCscPer];
    double sq = 0.0;
#pragma unroll
    for (int q = 0; q < kCscPer; ++q) {
        const int m = tid + q * kCscThreads;
        wr[q] = make_double2(0.0, 0.0);
        if (m < M) {
            const double rm = dist3(gx, gy, gz, a.mpos[m], a.mpos[M + m], a.mpos[2 * M + m]);
            float c, s;
            phasor(k * (rm - r0), a.phase_mode, c, s);
            const double ar = (double)c * r0 / rm, ai = -(double)s * r0 / rm;
            wr[q] = make_double2(ar, ai);
            sq += ar * ar + ai * ai;
        }
    }
    const double S = block_sum(sq, s_red, par); // sum |a|^2 (true level / location)
    const double sn = sqrt(a.norm);
#pragma unroll
    for (int q = 0; q < kCscPer; ++q) {
        const int m = tid + q * kCscThreads;
        if (m >= M) continue;
        const double2 x = wr[q];
        double vr, vi;
        switch (a.steer_type) {
        case ABF_STEER_CLASSIC: {
            const double ab = hypot(x.x, x.y);
            vr = x.x / ab / M;
            vi = x.y / ab / M;
            break;
        }
        case ABF_STEER_INVERSE: { // 1 / conj(x) / M
            const double d = x.x * x.x + x.y * x.y;
            vr = x.x / d / M;
            vi = x.y / d / M;
            break;
        }
        case ABF_STEER_TRUE_LEVEL:
            vr = x.x / S;
            vi = x.y / S;
            break;
        default: {
            const double d = sqrt(S * M);
            vr = x.x / d;
            vi = x.y / d;
            break;
        }
        }
        // wmax = conj(steer * sqrt(normfactor))
        wr[q] = make_double2(vr * sn, -(vi * sn));
        s_w[m] = wr[q];
    }
    __syncthreads();
    // 4. D1 = (C^T - diag C) w / hmax: D1_a = sum_{b != a} C[b][a] w_b / hmax.  Up to 256
    //    microphones the rows are split over kCscThreads / M groups of threads, whose partial
    //    sums are added in group order; the loads go out kCscBatch at a time.
    const double2 *C = a.csm + (int64_t)f * M * M;
    double2 d1[kCscPer], hr[kCscPer];
#pragma unroll
    for (int q = 0; q < kCscPer; ++q) {
        d1[q] = make_double2(0.0, 0.0);
        hr[q] = wr[q];
    }
    if (M <= kCscThreads) {
        const int groups = kCscThreads / M, am = tid % M, grp = tid / M;
        double dr = 0.0, di = 0.0;
        if (grp < groups) {
            for (int b0 = grp; b0 < M; b0 += kCscBatch * groups) {
                double2 c[kCscBatch];
#pragma unroll
                for (int u = 0; u < kCscBatch; ++u) {
                    const int b = b0 + u * groups;
                    c[u] = C[(int64_t)(b < M ? b : M - 1) * M + am];
                }
#pragma unroll
                for (int u = 0; u < kCscBatch; ++u) {
                    const int b = b0 + u * groups;
                    if (b < M && b != am) {
                        const double2 w = s_w[b];
                        dr += c[u].x * w.x - c[u].y * w.y;
                        di += c[u].x * w.y + c[u].y * w.x;
                    }
                }
            }
            s_part[tid] = make_double2(dr, di);
        }
        __syncthreads();
        if (tid < M) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < groups; ++q) {
                const double2 p = s_part[q * M + tid];
                sr += p.x;
                si += p.y;
            }
            d1[0] = make_double2(sr / hmax, si / hmax);
        }
    } else {
#pragma unroll
        for (int q = 0; q < kCscPer; ++q) {
            const int am = tid + q * kCscThreads;
            if (am >= M) continue;
            double dr = 0.0, di = 0.0;
            for (int b0 = 0; b0 < M; b0 += kCscBatch) {
                double2 c[kCscBatch];
#pragma unroll
                for (int u = 0; u < kCscBatch; ++u) {
                    const int b = b0 + u;
                    c[u] = C[(int64_t)(b < M ? b : M - 1) * M + am];
                }
#pragma unroll
                for (int u = 0; u < kCscBatch; ++u) {
                    const int b = b0 + u;
                    if (b < M && b != am) {
                        const double2 w = s_w[b];
                        dr += c[u].x * w.x - c[u].y * w.y;
                        di += c[u].x * w.y + c[u].y * w.x;
                    }
                }
            }
            d1[q] = make_double2(dr / hmax, di / hmax);
        }
    }
    // 5. hh <- (D1 + |hh|^2 w) / sqrt(1 + sum |w|^2 |hh|^2), 20 times; each thread keeps its own
    //    microphones in registers, only the sum is shared
    for (int it = 0; it < 20; ++it) {
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < kCscPer; ++q) {
            const double2 w = wr[q], x = hr[q];
            part += (w.x * w.x + w.y * w.y) * (x.x * x.x + x.y * x.y);
        }
        const double den = sqrt(1.0 + block_sum(part, s_red, par));
#pragma unroll
        for (int q = 0; q < kCscPer; ++q) {
            const double2 w = wr[q], x = hr[q], d = d1[q];
            const double H = x.x * x.x + x.y * x.y;
            hr[q] = make_double2((d.x + H * w.x) / den, (d.y + H * w.y) / den);
        }
    }
    // 6. hand hh and hmax to the map kernel; C -= damp * (hmax hh hh^H)^T
#pragma unroll
    for (int q = 0; q < kCscPer; ++q) {
        const int m = tid + q * kCscThreads;
        if (m < M) {
            s_hh[m] = hr[q];
            a.hh[(int64_t)f * M + m] = hr[q];
        }
    }
    if (tid == 0) a.hmax[f] = hmax;
    __syncthreads();
    double2 *Cw = a.csm + (int64_t)f * M * M;
    const int64_t MM = (int64_t)M * M;
    for (int64_t base = tid; base < MM; base += kCscBatch * kCscThreads) {
        double2 v[kCscBatch];
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t idx = base + (int64_t)u * kCscThreads;
            v[u] = Cw[idx < MM ? idx : MM - 1];
        }
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t idx = base + (int64_t)u * kCscThreads;
            if (idx >= MM) continue;
            const int r = (int)(idx / M), cidx = (int)(idx % M);
            const double2 hb = s_hh[cidx], ha = s_hh[r];
            // (hmax hh hh^H)[b][a] = hmax hh_b conj(hh_a)
            const double pr = hb.x * ha.x + hb.y * ha.y, pi = hb.y * ha.x - hb.x * ha.y;
            v[u].x -= a.damp * (hmax * pr);
            v[u].y -= a.damp * (hmax * pi);
            Cw[idx] = v[u];
        }
    }
}

__global__ __launch_bounds__(256) void csc_map_kernel(CscArgs a)
{
    __shared__ double2 s_hh[kCscMaxM];
    const int f = blockIdx.y;
    if (a.stopped[f]) return;
    const int M = a.M;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s_hh[m] = a.hh[(int64_t)f * M + m];
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.G) return;
    const double k = a.wavenumber[f];
    double tr = 0.0, ti = 0.0, ds = 0.0;
    const double2 *rw = a.rw + g;
    for (int m0 = 0; m0 < M; m0 += kCscBatch) {
        double2 t[kCscBatch];
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int m = m0 + u;
            t[u] = rw[(int64_t)(m < M ? m : M - 1) * a.G];
        }
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int m = m0 + u; // uniform over the workgroup
            if (m >= M) break;
            float c, s;
            phasor(k * t[u].x, a.phase_mode, c, s);
            const double vr = t[u].y * c, vi = -t[u].y * s;
            // conj(E_m) v_m with E = conj(hh): hh_m v_m
            const double2 q = s_hh[m];
            const double pr = q.x * vr - q.y * vi, pi = q.x * vi + q.y * vr;
            tr += pr;
            ti += pi;
            ds += (q.x * q.x + q.y * q.y) * (vr * vr + vi * vi);
        }
    }
    double full = tr * tr + ti * ti;
    if (a.r_diag) full -= ds;
    const double q = full * a.hmax[f];
    const double gx = a.gpos[g], gy = a.gpos[a.G + g], gz = a.gpos[2 * a.G + g];
    double res, nst;
    normalise(a.steer_type, q, csc_r0(a, gx, gy, gz), a.help[g], (double)M, a.norm, res, nst);
    a.h[(int64_t)f * a.G + g] -= a.damp * res;
}

struct CleanArgs {
    const double *gpos;       // [3][G]
    const double *mpos;       // [3][M]
    const double *wavenumber; // [F]
    double *dirty;            // [F][G] dirty maps, updated
    double *clean;            // [F][G] clean maps
    double *prev_sum;         // [F] sum |dirty| at the start of the last iteration
    double *amp;              // [F] amplitude of the last iteration
    int64_t *peak;            // [F] peak of the last iteration
    int *stopped;             // [F]
    const double2 *rw;        // [M][G] distance grid point -> mic, steering weight
    const double *help;       // [G] sum over the mics of 1 / r^2
    double ref[3];
    double ref_dist;
    double damp;
    int64_t G;
    int M, F, steer_type, phase_mode, psf_f32;
};

__device__ __forceinline__ double clean_r0(const CleanArgs &a, int64_t g)
{
    return a.ref_dist > 0.0 ? a.ref_dist
                            : dist3(a.gpos[g], a.gpos[a.G + g], a.gpos[2 * a.G + g], a.ref[0], a.ref[1], a.ref[2]);
}

__global__ __launch_bounds__(kCscThreads) void clean_select_kernel(CleanArgs a, int j)
{
    __shared__ double s_pv[4], s_ps[4];
    __shared__ int64_t s_pi[4];
    const int f = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t G = a.G;
    if (a.stopped[f]) return;
    // sum of absolute values and the peak (first index of the maximum) of the dirty map
    const double *d = a.dirty + (int64_t)f * G;
    double best = -INFINITY, asum = 0.0;
    int64_t bi = G;
    for (int64_t base = tid; base < G; base += kCscBatch * kCscThreads) {
        double v[kCscBatch];
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t g = base + (int64_t)u * kCscThreads;
            v[u] = d[g < G ? g : G - 1];
        }
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t g = base + (int64_t)u * kCscThreads;
            if (g < G) {
                asum += fabs(v[u]);
                if (v[u] > best) {
                    best = v[u];
                    bi = g;
                }
            }
        }
    }
    asum = __ockl_wfred_add_f64(asum);
    for (int off = 32; off > 0; off >>= 1) {
        const double v = __shfl_xor(best, off);
        const int64_t i = __shfl_xor(bi, off);
        if (v > best || (v == best && i < bi)) {
            best = v;
            bi = i;
        }
    }
    if (lane == 0) {
        s_pv[wv] = best;
        s_pi[wv] = bi;
        s_ps[wv] = asum;
    }
    __syncthreads();
    if (tid != 0) return;
    double dmax = s_pv[0];
    int64_t xi = s_pi[0];
    for (int q = 1; q < 4; ++q) {
        if (s_pv[q] > dmax || (s_pv[q] == dmax && s_pi[q] < xi)) {
            dmax = s_pv[q];
            xi = s_pi[q];
        }
    }
    if (xi >= G) xi = 0;
    const double S = ((s_ps[0] + s_ps[1]) + s_ps[2]) + s_ps[3];
    // the reference tests, after each iteration, that it lowered sum |dirty| and left a positive
    // value; the first iteration always runs
    if (j > 0 && !(a.prev_sum[f] > S && dmax > 0.0)) {
        a.stopped[f] = 1;
        return;
    }
    const double amp = a.damp * d[xi];
    a.clean[(int64_t)f * G + xi] += amp;
    a.amp[f] = amp;
    a.peak[f] = xi;
    a.prev_sum[f] = S;
}

__global__ __launch_bounds__(256) void clean_update_kernel(CleanArgs a)
{
    const int f = blockIdx.y;
    if (a.stopped[f]) return;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.G) return;
    const int M = a.M;
    const int64_t s = a.peak[f];
    const double k = a.wavenumber[f];
    double sr = 0.0, si = 0.0;
    const double2 *rg = a.rw + g;
    const double2 *rsrc = a.rw + s;
    for (int m0 = 0; m0 < M; m0 += kCscBatch) {
        double2 t[kCscBatch];
        double rs[kCscBatch];
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            const int64_t m = m0 + u < M ? m0 + u : M - 1;
            t[u] = rg[m * a.G];
            rs[u] = rsrc[m * a.G].x;
        }
#pragma unroll
        for (int u = 0; u < kCscBatch; ++u) {
            if (m0 + u >= M) break; // uniform over the workgroup
            float c, sn;
            phasor(k * (t[u].x - rs[u]), a.phase_mode, c, sn);
            // (cos - 1j sin) / r_s, times r_g (inverse) or 1 / r_g (true level, true location)
            double er = (double)c / rs[u], ei = -(double)sn / rs[u];
            if (a.steer_type != ABF_STEER_CLASSIC) {
                er *= t[u].y;
                ei *= t[u].y;
            }
            sr += er;
            si += ei;
        }
    }
    const double sq = sr * sr + si * si;
    const double r0s = clean_r0(a, s);
    double res;
    switch (a.steer_type) {
    case ABF_STEER_CLASSIC: {
        const double nf = r0s / M;
        res = sq * (nf * nf);
        break;
    }
    case ABF_STEER_INVERSE: {
        const double nf = r0s / clean_r0(a, g) / M;
        res = sq * (nf * nf);
        break;
    }
    case ABF_STEER_TRUE_LEVEL: {
        const double nf = r0s / clean_r0(a, g) / a.help[g];
        res = sq * (nf * nf);
        break;
    }
    default:
        res = sq * (r0s * r0s) / M / a.help[g];
        break;
    }
    if (a.psf_f32) res = (double)(float)res;
    a.dirty[(int64_t)f * a.G + g] -= res * a.amp[f];
}

__global__ __launch_bounds__(256) void fill_zero_f64_kernel(double *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Workspace layout: every array starts on a 256-byte boundary.
struct CscLayout {
    size_t powers, hh, hmax, stopped, rw, help, total;
};

inline CscLayout csc_layout(int M, int64_t G, int F, int J)
{
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    CscLayout l{};
    l.powers = 0;
    l.hh = up(l.powers + (size_t)F * J * sizeof(double));
    l.hmax = up(l.hh + (size_t)F * M * sizeof(double2));
    l.stopped = up(l.hmax + (size_t)F * sizeof(double));
    l.rw = up(l.stopped + (size_t)F * sizeof(int));
    l.help = up(l.rw + (size_t)M * G * sizeof(double2));
    l.total = up(l.help + (size_t)G * sizeof(double));
    return l;
}

struct CleanLayout {
    size_t prev_sum, amp, peak, stopped, rw, help, total;
};

inline CleanLayout clean_layout(int M, int64_t G, int F)
{
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    CleanLayout l{};
    l.prev_sum = 0;
    l.amp = up(l.prev_sum + (size_t)F * sizeof(double));
    l.peak = up(l.amp + (size_t)F * sizeof(double));
    l.stopped = up(l.peak + (size_t)F * sizeof(int64_t));
    l.rw = up(l.stopped + (size_t)F * sizeof(int));
    l.help = up(l.rw + (size_t)M * G * sizeof(double2));
    l.total = up(l.help + (size_t)G * sizeof(double));
    return l;
}

} // namespace abf

using namespace abf;

extern "C" {

size_t abf_cleansc_workspace_size(int32_t num_mics, int64_t num_grid, int32_t num_freqs, int32_t n_iter)
{
    if (num_mics < 1 || num_grid < 0 || num_freqs < 0 || n_iter < 0) return 0;
    return csc_layout(num_mics, num_grid, num_freqs, n_iter ? n_iter : 2 * num_mics).total;
}

int abf_cleansc(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, void *csm, double *h,
                int32_t steer_type, int32_t r_diag, double norm_factor, int32_t phase_mode, int32_t n_iter,
                double damp, int32_t stopn, float *result, void *workspace, size_t workspace_bytes, void *stream)
{
    if (!geom) return set_error(ABF_ERR_INVALID, "null geometry");
    if (steer_type < ABF_STEER_CLASSIC || steer_type > ABF_STEER_TRUE_LOCATION)
        return set_error(ABF_ERR_INVALID, "unknown steer_type");
    const int pm = phase_mode & 0xff;
    if (pm != ABF_PHASE_REFERENCE && pm != ABF_PHASE_FP64) return set_error(ABF_ERR_INVALID, "unknown phase_mode");
    const int M = geom->num_mics;
    const int64_t G = geom->num_grid;
    if (M < 1 || G < 0 || num_freqs < 0 || n_iter < 0) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (M > kCscMaxM) return set_error(ABF_ERR_UNSUPPORTED, "CLEAN-SC supports at most 1024 microphones");
    if (num_freqs > 65535) return set_error(ABF_ERR_INVALID, "at most 65535 frequencies per call");
    if (G == 0 || num_freqs == 0) return ABF_OK;
    if (!geom->gpos || !geom->mpos || !wavenumber || !csm || !h || !result)
        return set_error(ABF_ERR_INVALID, "null pointer");
    const size_t need = abf_cleansc_workspace_size(M, G, num_freqs, n_iter);
    if (!workspace || workspace_bytes < need) return set_error(ABF_ERR_WORKSPACE, "workspace too small");
    const int J = n_iter ? n_iter : 2 * M;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    CscArgs a{};
    a.gpos = geom->gpos;
    a.mpos = geom->mpos;
    a.wavenumber = wavenumber;
    a.csm = static_cast<double2 *>(csm);
    a.h = h;
    a.result = result;
    const CscLayout lay = csc_layout(M, G, num_freqs, J);
    a.powers = reinterpret_cast<double *>(ws + lay.powers);
    a.hh = reinterpret_cast<double2 *>(ws + lay.hh);
    a.hmax = reinterpret_cast<double *>(ws + lay.hmax);
    a.stopped = reinterpret_cast<int *>(ws + lay.stopped);
    a.rw = reinterpret_cast<double2 *>(ws + lay.rw);
    a.help = reinterpret_cast<double *>(ws + lay.help);
    for (int i = 0; i < 3; ++i) a.ref[i] = geom->ref[i];
    a.ref_dist = geom->ref_dist;
    a.norm = norm_factor;
    a.damp = damp;
    a.G = G;
    a.M = M;
    a.F = num_freqs;
    a.J = J;
    a.stopn = stopn;
    a.steer_type = steer_type;
    a.r_diag = r_diag;
    a.phase_mode = pm;
    // the result maps and the stop flags start at zero (fill kernels, no memset nodes)
    hipLaunchKernelGGL(fill_zero_f32_kernel, dim3((unsigned)((G * num_freqs + 255) / 256)), dim3(256), 0, s, result,
                       G * num_freqs);
    hipLaunchKernelGGL(fill_zero_i32_kernel, dim3((unsigned)((num_freqs + 255) / 256)), dim3(256), 0, s, a.stopped,
                       (int64_t)num_freqs);
    hipLaunchKernelGGL(csc_table_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, a.gpos, a.mpos, a.rw,
                       a.help, G, M, steer_type);
    for (int j = 0; j < J; ++j) {
        hipLaunchKernelGGL(csc_select_kernel, dim3((unsigned)num_freqs), dim3(kCscThreads), 0, s, a, j);
        hipLaunchKernelGGL(csc_map_kernel, dim3((unsigned)((G + 255) / 256), (unsigned)num_freqs), dim3(256), 0, s, a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

size_t abf_clean_workspace_size(int32_t num_mics, int64_t num_grid, int32_t num_freqs)
{
    if (num_mics < 1 || num_grid < 0 || num_freqs < 0) return 0;
    return clean_layout(num_mics, num_grid, num_freqs).total;
}

int abf_clean(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, double *dirty,
              int32_t steer_type, int32_t phase_mode, int32_t psf_f32, int32_t n_iter, double damp, double *clean,
              void *workspace, size_t workspace_bytes, void *stream)
{
    if (!geom) return set_error(ABF_ERR_INVALID, "null geometry");
    if (steer_type < ABF_STEER_CLASSIC || steer_type > ABF_STEER_TRUE_LOCATION)
        return set_error(ABF_ERR_INVALID, "unknown steer_type");
    const int pm = phase_mode & 0xff;
    if (pm != ABF_PHASE_REFERENCE && pm != ABF_PHASE_FP64) return set_error(ABF_ERR_INVALID, "unknown phase_mode");
    const int M = geom->num_mics;
    const int64_t G = geom->num_grid;
    if (M < 1 || G < 0 || num_freqs < 0) return set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_freqs > 65535) return set_error(ABF_ERR_INVALID, "at most 65535 frequencies per call");
    if (G == 0 || num_freqs == 0) return ABF_OK;
    if (!geom->gpos || !geom->mpos || !wavenumber || !dirty || !clean)
        return set_error(ABF_ERR_INVALID, "null pointer");
    const size_t need = abf_clean_workspace_size(M, G, num_freqs);
    if (!workspace || workspace_bytes < need) return set_error(ABF_ERR_WORKSPACE, "workspace too small");
    // the reference's loop runs its first iteration unconditionally
    const int J = n_iter > 1 ? n_iter : 1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    const CleanLayout lay = clean_layout(M, G, num_freqs);
    CleanArgs a{};
    a.gpos = geom->gpos;
    a.mpos = geom->mpos;
    a.wavenumber = wavenumber;
    a.dirty = dirty;
    a.clean = clean;
    a.prev_sum = reinterpret_cast<double *>(ws + lay.prev_sum);
    a.amp = reinterpret_cast<double *>(ws + lay.amp);
    a.peak = reinterpret_cast<int64_t *>(ws + lay.peak);
    a.stopped = reinterpret_cast<int *>(ws + lay.stopped);
    double2 *rw = reinterpret_cast<double2 *>(ws + lay.rw);
    double *help = reinterpret_cast<double *>(ws + lay.help);
    a.rw = rw;
    a.help = help;
    for (int i = 0; i < 3; ++i) a.ref[i] = geom->ref[i];
    a.ref_dist = geom->ref_dist;
    a.damp = damp;
    a.G = G;
    a.M = M;
    a.F = num_freqs;
    a.steer_type = steer_type;
    a.phase_mode = pm;
    a.psf_f32 = psf_f32 ? 1 : 0;
    // the clean maps and the stop flags start at zero (fill kernels, no memset nodes)
    hipLaunchKernelGGL(fill_zero_f64_kernel, dim3((unsigned)((G * num_freqs + 255) / 256)), dim3(256), 0, s, clean,
                       G * num_freqs);
    hipLaunchKernelGGL(fill_zero_i32_kernel, dim3((unsigned)((num_freqs + 255) / 256)), dim3(256), 0, s, a.stopped,
                       (int64_t)num_freqs);
    // CLEAN's PSF weights are those of the beamformer: r (inverse), 1/r (true level, location)
    hipLaunchKernelGGL(csc_table_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, a.gpos, a.mpos, rw,
                       help, G, M, steer_type);
    for (int j = 0; j < J; ++j) {
        hipLaunchKernelGGL(clean_select_kernel, dim3((unsigned)num_freqs), dim3(kCscThreads), 0, s, a, j);
        hipLaunchKernelGGL(clean_update_kernel, dim3((unsigned)((G + 255) / 256), (unsigned)num_freqs), dim3(256), 0,
                           s, a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
