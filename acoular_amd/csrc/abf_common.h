// Shared device helpers for the gfx950 beamforming kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/abf.h"

namespace abf {

// 2*pi split into a double head and the residual 2*pi - head (Cody-Waite reduction with a
// two-term constant: Cody & Waite, "Software Manual for the Elementary Functions", 1980).
constexpr double kTwoPiHi = 6.283185307179586232;
constexpr double kTwoPiLo = 2.4492935982947064e-16;
constexpr double kInvTwoPi = 0.15915494309189533577;

// cos/sin of a phase given in radians as a double.  With ABF_PHASE_REFERENCE the phase is
// first rounded to float32, as the reference does before it calls cos/sin
// (acoular/fastFuncs.py:260).  The argument is then reduced to [-pi, pi] in float64 and the
// float32 sincos is taken of the reduced value, so the error stays at float32 rounding of a
// number no larger than pi whatever the size of k*r.
__device__ __forceinline__ void phasor(double phi, int phase_mode, float &c, float &s)
{
    if (phase_mode == ABF_PHASE_REFERENCE) phi = (double)(float)phi;
    const double n = rint(phi * kInvTwoPi);
    double t = fma(-n, kTwoPiHi, phi);
    t = fma(-n, kTwoPiLo, t);
    sincosf((float)t, &s, &c);
}

// cos/sin of phi as phasor()'s, with fast != 0 taken by the hardware transcendental unit:
// v_sin/v_cos take their argument in revolutions, so the phase is reduced to [-1/2, 1/2]
// revolutions in float64 first.
__device__ __forceinline__ void phasor_rev(double phi, int phase_mode, int fast, float &c, float &s)
{
    if (phase_mode == ABF_PHASE_REFERENCE) phi = (double)(float)phi;
    if (fast) {
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

// round-to-nearest-even float -> bf16 bit pattern (finite inputs)
__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ float bf16_to_float(uint32_t b) { return __uint_as_float(b << 16); }

// split x = hi + lo with hi, lo bf16 (hi = RNE(x), lo = RNE(x - hi))
__device__ __forceinline__ void split_bf16(float x, uint32_t &hi, uint32_t &lo)
{
    hi = bf16_bits(x);
    lo = bf16_bits(x - bf16_to_float(hi));
}

__device__ __forceinline__ double dist3(double gx, double gy, double gz, double mx, double my, double mz)
{
    const double dx = gx - mx, dy = gy - my, dz = gz - mz;
    return sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
}

// Steering weight of one microphone (acoular/fastFuncs.py:262-273): inverse multiplies the
// phasor by r, true level/location divide by r; the sum feeding the normalisation is
// sum r^2 or sum 1/r^2 respectively.
__device__ __forceinline__ void steer_weight(int steer_type, double r, float &w, double &help)
{
    if (steer_type == ABF_STEER_INVERSE) {
        w = (float)r;
        help = r * r;
    } else if (steer_type == ABF_STEER_TRUE_LEVEL || steer_type == ABF_STEER_TRUE_LOCATION) {
        w = (float)(1.0 / r);
        help = 1.0 / (r * r);
    } else {
        w = 1.0f;
        help = 0.0;
    }
}

// Steer-type normalisation of the quadratic form (acoular/fastFuncs.py:293-310).
__device__ __forceinline__ void normalise(int steer_type, double q, double r0, double help, double nmics,
                                          double norm, double &result, double &nsteer)
{
    switch (steer_type) {
    case ABF_STEER_CLASSIC:
        result = q / (nmics * nmics) * norm;
        nsteer = 1.0 / nmics;
        break;
    case ABF_STEER_INVERSE: {
        const double nf = nmics * r0;
        result = q / (nf * nf) * norm;
        nsteer = help / (nf * nf);
        break;
    }
    case ABF_STEER_TRUE_LEVEL: {
        const double nf = r0 * help;
        result = q / (nf * nf) * norm;
        nsteer = 1.0 / (r0 * r0) / help;
        break;
    }
    default: {
        const double nf = nmics * help;
        result = q / nf * norm;
        nsteer = 1.0 / nmics;
        break;
    }
    }
}

// K-slot order inside a 32-microphone block of the MFMA forms: slot (g, e) -- lane group g =
// lane >> 4, element e of the lane's 8-element fragment -- holds microphone
// 4 g + (e & 3) + 16 (e >> 2).  With it, the B fragment of lane (column, g) holds exactly the
// microphones 16 t + 4 g + i (t = e >> 2, i = e & 3) that the 16x16 MFMA result rows 4 g + i of
// row tile t of the same block hold, so the epilogue q = Re sum conj(x) Y needs no lane exchange.
__device__ __forceinline__ int kslot_mic(int g, int e) { return 4 * g + (e & 3) + 16 * (e >> 2); }

__host__ __device__ __forceinline__ int weight_kind(int steer_type)
{
    return steer_type == ABF_STEER_CLASSIC ? 0 : (steer_type == ABF_STEER_INVERSE ? 1 : 2);
}

} // namespace abf
