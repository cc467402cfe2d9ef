// Kernel argument blocks and launch declarations shared between the kernel sources and the
// C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abf {

struct PlainArgs {
    const float2 *ut;          // [F][Mp][Mp] U^T
    const double *gpos;        // [3][G] or null when rm_table is used
    const double *mpos;        // [3][M]
    const double *rm_table;    // [G][M] or null
    const double *r0_table;    // [G] or null
    const double *wavenumber;  // [F]
    float *out;                // [F][G]
    float *out_norm;           // [F][G] or null
    double ref[3];
    double ref_dist;
    double norm;
    int64_t G;
    int M, Mp;
    int steer_type, phase_mode, clip, fast_trig;
};

// float64 plain form (U^T and steering vectors in float64, float64 outputs)
struct Plain64Args {
    const double2 *ut;         // [F][Mp][Mp] U^T
    const double *gpos;        // [3][G]
    const double *mpos;        // [3][M]
    const double *wavenumber;  // [F]
    double *out;               // [F][G]
    double *out_norm;          // [F][G] or null
    double ref[3];
    double ref_dist;
    double norm;
    int64_t G;
    int M, Mp;
    int steer_type, phase_mode, clip;
};

// one beamforming job of an MFMA launch
// Operand precision of the MFMA forms (template parameter PR of the kernels)
constexpr int kPrF32 = 0;    // fp32 MFMA, U and X in float32
constexpr int kPrBF16X3 = 1; // bf16 MFMA, U and X split hi + lo, 3 products per real product
constexpr int kPrF16X2 = 2;  // fp16 MFMA, U rounded once (scaled per frequency), X split hi + lo: 2 products

struct MfJob {
    const void *prep;          // U in fragment order, [F][nUpper][2][4 KiB]
    const double *uscale;      // [F] power-of-two factor undoing the fp16 scaling of U (f16x2 form), or null
    const double *gpos;        // [3][G]
    const double *mpos;        // [3][M]
    const double *wavenumber;  // [F]
    float *out;                // [F][G]
    float *out_norm;           // [F][G] or null
    double ref[3];
    double ref_dist;
    double norm;
    int64_t G;
    int64_t wg_begin;          // first workgroup of this job in the launch
    int M, Mp, CT, F, FB, ntiles;
    int steer_type, phase_mode, clip, fast_trig;
};

struct EigArgs {
    const double *eigval;      // [K]
    const float2 *eigvec_t;    // [K][M] (transposed eigenvectors)
    const float2 *steer;       // [G][M] explicit steering or null
    const double *r0_table;    // [G]
    const double *rm_table;    // [G][M]
    float *out, *out_norm;
    double wavenumber, norm;
    int64_t G;
    int M, K, steer_type, r_diag, phase_mode, fast_trig;
};

// eigenvalue form over every frequency of a job, steering from coordinates
struct EigGridArgs {
    const double *gpos;        // [3][G]
    const double *mpos;        // [3][M]
    const double *wavenumber;  // [F]
    const double *eigval;      // [F][M]
    const float2 *eigvec;      // [F][M][M], element (m, e) = component m of eigenvector e
    float *out;                // [F][G]
    float *out_norm;           // [F][G] or null
    double ref[3];
    double ref_dist;
    double norm;
    int64_t G;
    int M, e0, K, P;           // eigenvectors e0 .. e0+K-1; P grid points per workgroup
    int steer_type, r_diag, phase_mode, fast_trig, clip;
};

struct CustomArgs {
    const float2 *csm;         // [M][M]
    const float2 *steer;       // [G][M]
    float *out, *out_norm;
    double norm;
    int64_t G;
    int M, r_diag;
};

__global__ void prep_plain_kernel(const float2 *csm, int M, int Mp, int r_diag, float2 *ut);
// microphones per LDS chunk of the plain kernel unless ABF_PLAIN_JC says otherwise
constexpr int kPlainDefaultJC = 64;
// LU: U staged in LDS (read by every lane) instead of broadcast through scalar loads
template <int JC, bool LU> __global__ void bf_plain_kernel(PlainArgs a);
__global__ void prep_plain64_kernel(const double2 *csm, int M, int Mp, int r_diag, double2 *ut);
__global__ void bf_plain64_kernel(Plain64Args a);
__global__ void prep_mfma_kernel(const float2 *csm, int M, int Mp, int r_diag, int prec, const double *uscale,
                                 char *prep);
__global__ void uscale_kernel(const float2 *csm, int M, int r_diag, double *uscale);
// VAR = fast trig (bit 0) | steering weight kind (0 none, 1 r, 2 1/r) << 1
// TT = row tiles (of 16) per work unit: 1, or 2 (both halves of a 32-row block); PF = software-
// pipelined X/U fragment loads (split-bf16 only)
template <int PR, int CW, int VAR, int TT, bool PF>
__global__ void bf_mfma_kernel(const MfJob *jobs, int njobs);
template <int PR, int VAR> __global__ void bf_mfma_ws_kernel(const MfJob *jobs, int njobs);

// streamed split-bf16 kernel (abf_mfma_stream.h): NK = Mp / 32 in 2..8, WK = steering weight kind
constexpr int kStWaves = 8;
constexpr int kStBN = 64;                // grid points per workgroup
constexpr int kStKSB = 40;               // LDS row stride of a steering block, bf16 elements (32 + 16-byte skew)
constexpr int kStPlane = kStBN * kStKSB; // elements per plane of a block
constexpr int kStBlock = 4 * kStPlane;   // elements per block (re hi, re lo, im hi, im lo)
#ifndef ABF_ST_GS
#define ABF_ST_GS 4
#endif
constexpr int kStGS = ABF_ST_GS;         // U blocks (pairs) per sub-step (3, 5: 2 % slower, 6 spills)
#ifndef ABF_ST_RING
#define ABF_ST_RING 3
#endif
// slots of the LDS ring of U sub-steps: the loads of sub-step t + kStRing - 1 are issued in sub-step t
// (3: 14.29 against 14.62 ms with 2 on the 256-mic, 201 x 201, 513-bin workload)
constexpr int kStRing = ABF_ST_RING;
static_assert(kStRing == 2 || kStRing == 3, "a U ring of two or three sub-steps");
// LDS layout: steering bf16 [2][4][BN][KSB], U ring [kStRing][kStGS][8 KiB], q partials [2 parities]
// [2 halves][BN] f32, normalisation factor [BN] f64, steering norm [BN] f32, r0 [BN] f64,
// steering-sum partials [2][BN] f64
constexpr size_t kStOffU = 2 * (size_t)kStBlock * 2;
constexpr size_t kStOffQ = kStOffU + kStRing * (size_t)kStGS * 8192;
constexpr size_t kStOffFac = kStOffQ + 4 * kStBN * 4;
constexpr size_t kStOffNst = kStOffFac + kStBN * 8;
constexpr size_t kStOffR0 = kStOffNst + kStBN * 4;
constexpr size_t kStOffHelp = kStOffR0 + kStBN * 8;
constexpr size_t st_lds_bytes() { return kStOffHelp + 2 * kStBN * 8; }
// PS: persistent launch (a grid of at most one workgroup per CU walking the work items)
template <int NK, int WK, bool PS> __global__ void bf_mfma_stream_kernel(const MfJob *jobs, int njobs, int64_t total);

// streamed split-bf16 kernel for up to 16 blocks of 32 microphones (abf_mfma_wide.h): 8 waves over
// 32 grid points, each one 16-column tile x one row half of the row blocks of one parity
constexpr int kWdWaves = 8;
constexpr int kWdBN = 32;                // grid points per workgroup
constexpr int kWdMaxNK = 16;             // blocks of 32 microphones (512 microphones)
constexpr int kWdKSB = 40;               // LDS row stride of a steering block, bf16 elements (32 + 16-byte skew)
constexpr int kWdPlane = kWdBN * kWdKSB; // elements per plane of a block
constexpr int kWdBlock = 4 * kWdPlane;   // elements per block (re hi, re lo, im hi, im lo)
constexpr int kWdGS = 4;                 // U blocks (pairs) per sub-step (5: 51.7 against 50.3 ms on the fused 512-mic class)
constexpr int kWdRing = kStRing;         // slots of the LDS ring of U sub-steps
// LDS layout: steering bf16 [2][4][BN][KSB], U ring [kWdRing][kWdGS][8 KiB], q partials [2 parities]
// [4 wave groups][BN] f32, normalisation factor [BN] f64, steering norm [BN] f32, r0 [BN] f64,
// steering-sum partials [4][BN] f64
constexpr size_t kWdOffU = 2 * (size_t)kWdBlock * 2;
constexpr size_t kWdOffQ = kWdOffU + kWdRing * (size_t)kWdGS * 8192;
constexpr size_t kWdOffFac = kWdOffQ + 2 * 4 * kWdBN * 4;
constexpr size_t kWdOffNst = kWdOffFac + kWdBN * 8;
constexpr size_t kWdOffR0 = kWdOffNst + kWdBN * 4;
constexpr size_t kWdOffHelp = kWdOffR0 + kWdBN * 8;
constexpr size_t wd_lds_bytes() { return kWdOffHelp + 4 * kWdBN * 8; }
template <int NK, int WK> __global__ void bf_mfma_wide_kernel(const MfJob *jobs, int njobs);
__global__ void bf_eig_kernel(EigArgs a);
__global__ void bf_eig_grid_kernel(EigGridArgs a);
__global__ void bf_eig_mfma_kernel(EigGridArgs a);
__global__ void bf_custom_kernel(CustomArgs a);

} // namespace abf
