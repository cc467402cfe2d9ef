// CLEAN-SC deconvolution (BeamformerCleansc._calc, acoular/fbeamform.py:1449-1492) for gfx950.
//
// The reference runs, per frequency, up to J iterations of: peak of the map h, coherent source
// component hh of the CSM at the peak (a 20-step fixed point on the steering vector), removal of
// hh's map from h and of hmax hh hh^H from the CSM.  Here every frequency of a request advances
// together, two launches per iteration:
//  * csc_select_kernel, one workgroup per frequency: the peak (first index of the maximum, as
//    numpy's argmax), the float32 result update, the stop test, the steering vector at the peak
//    (transfer function, SteeringVector.steer_vector normalisation, fbeamform.py:145-153,
//    207-258), D1 = (C^T - diag C) w / hmax, the fixed point, and the CSM update;
//  * csc_map_kernel, one thread per (grid point, frequency): h -= damp * (the eigenvalue-form map
//    of the single pair (hmax, conj(hh)), fastFuncs.py:335-411).
// A frequency whose stop test fires is skipped by all later launches, so the result equals the
// reference's early break.  Arithmetic is float64 except the reference's float32 phases and
// float32 cos/sin (fastFuncs.py:259-261, 896-897).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "abf_common.h"
#include "abf_status.h"

namespace abf {

constexpr int kCscMaxM = 1024;

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

__device__ double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    return red[0];
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

__global__ __launch_bounds__(256) void csc_select_kernel(CscArgs a, int j)
{
    __shared__ double s_val[256];
    __shared__ int64_t s_idx[256];
    __shared__ double s_red[256];
    __shared__ double2 s_w[kCscMaxM], s_d1[kCscMaxM], s_hh[kCscMaxM];
    __shared__ int s_stop;
    const int f = blockIdx.x, tid = threadIdx.x;
    const int M = a.M;
    const int64_t G = a.G;
    if (a.stopped[f]) return;
    // 1. the peak of h: largest value, first index on ties
    const double *h = a.h + (int64_t)f * G;
    double best = -INFINITY;
    int64_t bi = G;
    for (int64_t g = tid; g < G; g += blockDim.x) {
        const double v = h[g];
        if (v > best) {
            best = v;
            bi = g;
        }
    }
    s_val[tid] = best;
    s_idx[tid] = bi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double v = s_val[tid + s];
            const int64_t i = s_idx[tid + s];
            if (v > s_val[tid] || (v == s_val[tid] && i < s_idx[tid])) {
                s_val[tid] = v;
                s_idx[tid] = i;
            }
        }
        __syncthreads();
    }
    const double hmax = s_val[0];
    const int64_t xi = s_idx[0] < G ? s_idx[0] : 0;
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
    double sq = 0.0;
    for (int m = tid; m < M; m += blockDim.x) {
        const double rm = dist3(gx, gy, gz, a.mpos[m], a.mpos[M + m], a.mpos[2 * M + m]);
        float c, s;
        phasor(k * (rm - r0), a.phase_mode, c, s);
        const double ar = (double)c * r0 / rm, ai = -(double)s * r0 / rm;
        s_w[m] = make_double2(ar, ai);
        sq += ar * ar + ai * ai;
    }
    const double S = block_sum(sq, s_red); // sum |a|^2 (true level / location)
    const double sn = sqrt(a.norm);
    for (int m = tid; m < M; m += blockDim.x) {
        const double2 x = s_w[m];
        double wr, wi;
        switch (a.steer_type) {
        case ABF_STEER_CLASSIC: {
            const double ab = hypot(x.x, x.y);
            wr = x.x / ab / M;
            wi = x.y / ab / M;
            break;
        }
        case ABF_STEER_INVERSE: { // 1 / conj(x) / M
            const double d = x.x * x.x + x.y * x.y;
            wr = x.x / d / M;
            wi = x.y / d / M;
            break;
        }
        case ABF_STEER_TRUE_LEVEL:
            wr = x.x / S;
            wi = x.y / S;
            break;
        default: {
            const double d = sqrt(S * M);
            wr = x.x / d;
            wi = x.y / d;
            break;
        }
        }
        // wmax = conj(steer * sqrt(normfactor))
        s_w[m] = make_double2(wr * sn, -(wi * sn));
    }
    __syncthreads();
    // 4. D1 = (C^T - diag C) w / hmax: D1_a = sum_{b != a} C[b][a] w_b / hmax
    const double2 *C = a.csm + (int64_t)f * M * M;
    for (int am = tid; am < M; am += blockDim.x) {
        double dr = 0.0, di = 0.0;
        for (int b = 0; b < M; ++b) {
            if (b == am) continue;
            const double2 c = C[(int64_t)b * M + am];
            const double2 w = s_w[b];
            dr += c.x * w.x - c.y * w.y;
            di += c.x * w.y + c.y * w.x;
        }
        s_d1[am] = make_double2(dr / hmax, di / hmax);
        s_hh[am] = s_w[am];
    }
    // 5. hh <- (D1 + |hh|^2 w) / sqrt(1 + sum |w|^2 |hh|^2), 20 times
    for (int it = 0; it < 20; ++it) {
        double part = 0.0;
        __syncthreads();
        for (int m = tid; m < M; m += blockDim.x) {
            const double2 w = s_w[m], q = s_hh[m];
            part += (w.x * w.x + w.y * w.y) * (q.x * q.x + q.y * q.y);
        }
        const double den = sqrt(1.0 + block_sum(part, s_red));
        for (int m = tid; m < M; m += blockDim.x) {
            const double2 w = s_w[m], q = s_hh[m], d = s_d1[m];
            const double H = q.x * q.x + q.y * q.y;
            s_hh[m] = make_double2((d.x + H * w.x) / den, (d.y + H * w.y) / den);
        }
    }
    __syncthreads();
    // 6. hand hh and hmax to the map kernel; C -= damp * (hmax hh hh^H)^T
    for (int m = tid; m < M; m += blockDim.x) a.hh[(int64_t)f * M + m] = s_hh[m];
    if (tid == 0) a.hmax[f] = hmax;
    double2 *Cw = a.csm + (int64_t)f * M * M;
    for (int64_t idx = tid; idx < (int64_t)M * M; idx += blockDim.x) {
        const int r = (int)(idx / M), cidx = (int)(idx % M);
        const double2 hb = s_hh[cidx], ha = s_hh[r];
        // (hmax hh hh^H)[b][a] = hmax hh_b conj(hh_a)
        const double pr = hb.x * ha.x + hb.y * ha.y, pi = hb.y * ha.x - hb.x * ha.y;
        double2 v = Cw[idx];
        v.x -= a.damp * (hmax * pr);
        v.y -= a.damp * (hmax * pi);
        Cw[idx] = v;
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
    const double gx = a.gpos[g], gy = a.gpos[a.G + g], gz = a.gpos[2 * a.G + g];
    const double k = a.wavenumber[f];
    double tr = 0.0, ti = 0.0, ds = 0.0, help = 0.0;
    for (int m = 0; m < M; ++m) {
        const double r = dist3(gx, gy, gz, a.mpos[m], a.mpos[M + m], a.mpos[2 * M + m]);
        float c, s;
        phasor(k * r, a.phase_mode, c, s);
        double w = 1.0;
        if (a.steer_type == ABF_STEER_INVERSE) {
            w = r;
            help += r * r;
        } else if (a.steer_type != ABF_STEER_CLASSIC) {
            w = 1.0 / r;
            help += 1.0 / (r * r);
        }
        const double vr = w * c, vi = -w * s;
        // conj(E_m) v_m with E = conj(hh): hh_m v_m
        const double2 q = s_hh[m];
        const double pr = q.x * vr - q.y * vi, pi = q.x * vi + q.y * vr;
        tr += pr;
        ti += pi;
        ds += (q.x * q.x + q.y * q.y) * (vr * vr + vi * vi);
    }
    double full = tr * tr + ti * ti;
    if (a.r_diag) full -= ds;
    const double q = full * a.hmax[f];
    double res, nst;
    normalise(a.steer_type, q, csc_r0(a, gx, gy, gz), help, (double)M, a.norm, res, nst);
    a.h[(int64_t)f * a.G + g] -= a.damp * res;
}

} // namespace abf

using namespace abf;

extern "C" {

size_t abf_cleansc_workspace_size(int32_t num_mics, int32_t num_freqs, int32_t n_iter)
{
    if (num_mics < 1 || num_freqs < 0 || n_iter < 0) return 0;
    const int J = n_iter ? n_iter : 2 * num_mics;
    const size_t F = (size_t)num_freqs;
    return F * J * sizeof(double) + F * num_mics * sizeof(double2) + F * sizeof(double) + F * sizeof(int) + 1024;
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
    const size_t need = abf_cleansc_workspace_size(M, num_freqs, n_iter);
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
    a.powers = reinterpret_cast<double *>(ws);
    a.hh = reinterpret_cast<double2 *>(ws + (size_t)num_freqs * J * sizeof(double));
    a.hmax = reinterpret_cast<double *>(reinterpret_cast<char *>(a.hh) + (size_t)num_freqs * M * sizeof(double2));
    a.stopped = reinterpret_cast<int *>(reinterpret_cast<char *>(a.hmax) + (size_t)num_freqs * sizeof(double));
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
    for (int j = 0; j < J; ++j) {
        hipLaunchKernelGGL(csc_select_kernel, dim3((unsigned)num_freqs), dim3(256), 0, s, a, j);
        hipLaunchKernelGGL(csc_map_kernel, dim3((unsigned)((G + 255) / 256), (unsigned)num_freqs), dim3(256), 0, s, a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
