// Geometry and steering-vector kernels: dist_mat (acoular/environments.py:65-103) and
// calcTransfer + steer_vector (acoular/fastFuncs.py:859-897, acoular/fbeamform.py:145-153).
#include "abf_common.h"
#include "abf_steer_kernels.h"

namespace abf {

// grid: (groups of kDistPts points from g0 on) x ceil(M / 512) workgroups; workgroup (gp, mb)
// writes rm[g][512 mb ...] for the kDistPts points g of group gp, two microphones per thread (one
// 16-byte store when M is even, every row then starting on an even element).  The microphones'
// coordinates stay in registers across the points (the grid coordinates are wave-uniform); one
// division per workgroup, none per element.
__global__ __launch_bounds__(256) void dist_mat_kernel(const double *__restrict__ gpos, const double *__restrict__ mpos,
                                                       int64_t G, int M, int64_t g0, int vec2,
                                                       double *__restrict__ rm)
{
    const int nmb = (M + 511) / 512;
    const int64_t gb = g0 + (int64_t)(blockIdx.x / nmb) * kDistPts;
    const int m = (int)(blockIdx.x % nmb) * 512 + 2 * threadIdx.x;
    if (m >= M) return;
    const bool two = m + 1 < M;
    const double mx0 = mpos[m], my0 = mpos[M + m], mz0 = mpos[2 * M + m];
    const double mx1 = two ? mpos[m + 1] : 0.0, my1 = two ? mpos[M + m + 1] : 0.0, mz1 = two ? mpos[2 * M + m + 1] : 0.0;
    const int np = (int)min((int64_t)kDistPts, G - gb);
    for (int i = 0; i < np; ++i) {
        const int64_t g = gb + i;
        const double gx = gpos[g], gy = gpos[G + g], gz = gpos[2 * G + g];
        const double d0 = dist3(gx, gy, gz, mx0, my0, mz0);
        if (two) {
            const double d1 = dist3(gx, gy, gz, mx1, my1, mz1);
            if (vec2) {
                *reinterpret_cast<double2 *>(rm + g * M + m) = make_double2(d0, d1);
                continue;
            }
            rm[g * M + m + 1] = d1;
        }
        rm[g * M + m] = d0;
    }
}

// One wave per grid point, the lanes over the microphones (rows of rm read and of out written
// contiguously): a[g][m] = exp(-1j f32(k (rm - r0))) r0 / rm, then the optional steer_vector
// normalisation, sum |a|^2 by a wave reduction; float64 except for the float32 phase (as the
// reference).  grid: ceil(G / 4) workgroups of 4 waves.
__global__ __launch_bounds__(256) void transfer_kernel(const double *__restrict__ r0, const double *__restrict__ rm,
                                                       int64_t G, int M, double k, int phase_mode, int steer_mode,
                                                       int out_c128, void *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= G) return; // uniform per wave
    const double rr0 = r0[g];
    const double *row = rm + g * M;
    double s2 = 0.0; // sum |a|^2 for true level / true location
    if (steer_mode == 3 || steer_mode == 4) {
        for (int m = lane; m < M; m += 64) {
            const double q = rr0 / row[m];
            s2 += q * q;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o, 64);
    }
    for (int m = lane; m < M; m += 64) {
        const double r = row[m];
        float c, s;
        phasor(k * (r - rr0), phase_mode, c, s);
        const double amp = rr0 / r;
        double ar = (double)c * amp, ai = -(double)s * amp;
        switch (steer_mode) {
        case 1: { // x / |x| / M
            const double mag = sqrt(ar * ar + ai * ai);
            ar = ar / mag / M;
            ai = ai / mag / M;
            break;
        }
        case 2: { // 1 / conj(x) / M
            const double d = ar * ar + ai * ai;
            const double br = ar / d, bi = ai / d; // 1/conj(x) = x / |x|^2
            ar = br / M;
            ai = bi / M;
            break;
        }
        case 3:
            ar /= s2;
            ai /= s2;
            break;
        case 4: {
            const double d = sqrt(s2 * M);
            ar /= d;
            ai /= d;
            break;
        }
        default:
            break;
        }
        if (out_c128) {
            reinterpret_cast<double2 *>(out)[g * M + m] = make_double2(ar, ai);
        } else {
            reinterpret_cast<float2 *>(out)[g * M + m] = make_float2((float)ar, (float)ai);
        }
    }
}

} // namespace abf
