// C-ABI entry points of libabf.so (declared in include/abf.h).  Host-side planning,
// validation and launches; the kernels live in abf_bf.hip, abf_steer.hip and abf_csm.hip.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "abf_common.h"
#include "abf_csm_kernels.h"
#include "abf_kernels.h"
#include "abf_status.h"
#include "abf_steer_kernels.h"

using namespace abf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_status(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ABF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return ABF_OK;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }
inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr size_t kLdsMax = 163840; // gfx950 LDS per workgroup
constexpr int kPhaseMask = 0xff;
constexpr int kFastTrigFlag = 0x100;

int check_steer(int st)
{
    if (st < ABF_STEER_CLASSIC || st > ABF_STEER_TRUE_LOCATION)
        return fail(ABF_ERR_INVALID, "unknown steer_type " + std::to_string(st));
    return ABF_OK;
}

int check_phase(int pm)
{
    const int m = pm & kPhaseMask;
    if (m != ABF_PHASE_REFERENCE && m != ABF_PHASE_FP64) return fail(ABF_ERR_INVALID, "unknown phase_mode");
    return ABF_OK;
}

// ---------------------------------------------------------------------------------------------
// MFMA planning
// ---------------------------------------------------------------------------------------------
struct MfPlan {
    int Mp, CT, CW, FB, nK, nUp;
    int TT;  // row tiles per work unit (alternating kernel)
    bool pf; // software-pipelined fragment loads (split-bf16 alternating kernel)
    bool ws; // warp-specialised kernel (double-buffered steering matrix)
    bool st; // streamed kernel (steering blocks through an LDS ring)
    bool wide; // streamed kernel with two column tiles per wave
    bool ps;   // streamed kernel, persistent launch (one workgroup per CU walking the work items)
    int64_t ntiles, wgs;
    size_t prep_bytes, lds;
    size_t uscale_off; // f16x2: offset of the per-frequency U scales in the job's workspace
    int prec;          // kPrF32, kPrBF16X3 or kPrF16X2
};

size_t mfma_lds(bool h16, int BN, int Mp)
{
    // X planes of the tile, r0 + normalisation sums, per-wave partial results, weight scales
    const size_t KS = h16 ? Mp + 8 : Mp + 4;
    const size_t planes = h16 ? 4 * (size_t)BN * KS * 2 : 2 * (size_t)BN * KS * 4;
    return planes + 16 * (size_t)BN + 4 * 8 * (size_t)BN + 8 * (size_t)BN;
}

size_t mfma_ws_lds(bool bf16, int BN, int Mp)
{
    // two X buffers (8 bytes per element in both forms), q partials [2][8][BN], r0, steering-sum
    // partials [256 / BN][BN]
    const size_t KS = bf16 ? Mp + 8 : Mp + 4;
    return 16 * (size_t)BN * KS + 2 * 8 * 4 * (size_t)BN + 8 * (size_t)BN + 8 * 256;
}

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

int plan_mfma(const abf_grid_job *j, MfPlan &p)
{
    p.prec = j->form == ABF_FORM_MFMA_BF16X3 ? kPrBF16X3 : (j->form == ABF_FORM_MFMA_F16X2 ? kPrF16X2 : kPrF32);
    const bool h16 = p.prec != kPrF32;
    p.Mp = round_up(j->geom.num_mics, 32);
    p.nK = p.Mp / 32;
    p.nUp = p.nK * (p.nK + 1) / 2;
    const int fb = j->tile_freqs > 0 ? j->tile_freqs : env_int("ABF_MFMA_FB", 8);
    p.FB = std::max(1, std::min(fb, j->num_freqs));
    p.prep_bytes = (size_t)j->num_freqs * p.nUp * 8192;
    p.uscale_off = p.prep_bytes;
    if (p.prec == kPrF16X2) p.prep_bytes += align256((size_t)j->num_freqs * sizeof(double));
    // grid points needed by this job, as a power-of-two count of 16-point tiles
    int ct_cap = 16;
    const int64_t tiles16 = cdiv(j->geom.num_grid, 16);
    while (ct_cap > 1 && ct_cap / 2 >= tiles16) ct_cap /= 2;
    const int ct_req = j->tile_cols > 0 ? std::max(1, j->tile_cols / 16) : env_int("ABF_MFMA_CT", 0);
    // The kernel that alternates steering generation and MFMA work over one LDS buffer by
    // default (the faster one where both apply, measured on the 256-mic workload); the
    // warp-specialised kernel on request (tile_kernel, or ABF_MFMA_KERNEL=2) when its two
    // steering buffers fit LDS with 32 or 64 grid points per tile.
    int kreq = j->tile_kernel > 0 ? j->tile_kernel : env_int("ABF_MFMA_KERNEL", ABF_MFMA_KERNEL_AUTO);
    // the f16x2 form exists in the alternating kernels only: the warp-specialised and pipelined
    // requests map to their alternating counterparts
    if (p.prec == kPrF16X2) {
        if (kreq == ABF_MFMA_KERNEL_SPECIALISED || kreq == ABF_MFMA_KERNEL_PIPELINED) kreq = ABF_MFMA_KERNEL_ALTERNATING;
        if (kreq == ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS) kreq = ABF_MFMA_KERNEL_ROW_PAIRS;
    }
    p.CT = 0;
    p.ws = false;
    p.st = false;
    p.wide = false;
    p.ps = false;
    if (kreq == ABF_MFMA_KERNEL_STREAMED || kreq == ABF_MFMA_KERNEL_STREAMED_WIDE ||
        kreq == ABF_MFMA_KERNEL_STREAMED_PERSISTENT) {
        // split-bf16, 2..8 blocks of 32 microphones (2..16 for the wide kernel), the reference's
        // float32 phase with the hardware sin/cos; anything else runs the row-pair kernel
        const bool wide = kreq == ABF_MFMA_KERNEL_STREAMED_WIDE;
        if (p.prec == kPrBF16X3 && p.Mp >= 64 && p.Mp <= (wide ? 32 * kWdMaxNK : 256) &&
            (j->phase_mode & kFastTrigFlag) && (j->phase_mode & kPhaseMask) == ABF_PHASE_REFERENCE) {
            p.st = true;
            p.wide = wide;
            // the streamed kernels hold the wavenumbers of a band in the 64 lanes of a register
            p.FB = std::min(p.FB, 64);
            p.ps = kreq == ABF_MFMA_KERNEL_STREAMED_PERSISTENT;
            const int bn = wide ? kWdBN : kStBN;
            p.CT = bn / 16;
            p.TT = 1;
            p.CW = 1;
            p.pf = false;
            p.ntiles = cdiv(j->geom.num_grid, bn);
            p.wgs = p.ntiles * cdiv(j->num_freqs, p.FB);
            p.lds = p.wide ? wd_lds_bytes() : st_lds_bytes();
            return ABF_OK;
        }
        kreq = ABF_MFMA_KERNEL_ROW_PAIRS;
    }
    if (kreq == ABF_MFMA_KERNEL_SPECIALISED) {
        for (int ct = 4; ct >= 2; ct /= 2) { // tiles of 2 column groups of 32 grid points
            if (ct > ct_cap || (ct_req > 0 && ct > ct_req)) continue;
            const int BN = 16 * ct;
            if (p.Mp / 8 > (256 / BN) * 5) continue; // 4 steering waves, at most 5 groups of 8 per thread
            if (mfma_ws_lds(h16, BN, p.Mp) > kLdsMax) continue;
            // the per-wave step table of one frequency lives in the 64 lanes of a register
            if (cdiv(2 * p.nK * (ct / 2), 4) * p.nK > 64) continue;
            p.CT = ct;
            p.ws = true;
            break;
        }
    }
    if (p.CT == 0) {
        for (int ct = 16; ct >= 1; ct /= 2) {
            if (ct > ct_cap || (ct_req > 0 && ct > ct_req)) continue;
            const int BN = 16 * ct;
            const int TPC = 512 / BN;
            // steering prologue: at most 5 groups of 8 per thread (4 for the pipelined kernels)
            const bool pf_req = kreq == ABF_MFMA_KERNEL_PIPELINED || kreq == ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS;
            if (p.Mp / 8 > TPC * (pf_req && p.prec == kPrBF16X3 ? 4 : 5)) continue;
            if (mfma_lds(h16, BN, p.Mp) > kLdsMax) continue;
            p.CT = ct;
            break;
        }
    }
    if (p.CT == 0)
        return fail(ABF_ERR_UNSUPPORTED, "MFMA form supports at most ~1150 microphones; use the plain form");
    p.TT = (kreq == ABF_MFMA_KERNEL_ROW_PAIRS || kreq == ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS) ? 2 : 1;
    p.pf = !p.ws && p.prec == kPrBF16X3 &&
           (kreq == ABF_MFMA_KERNEL_PIPELINED || kreq == ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS);
    if (p.ws) p.CW = 2;
    else if (kreq == ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS) p.CW = 1;
    else if (p.TT == 2 || kreq == ABF_MFMA_KERNEL_PIPELINED) p.CW = std::min(p.CT, 2);
    else p.CW = std::min(p.CT, 4);
    p.ntiles = cdiv(j->geom.num_grid, 16 * p.CT);
    p.wgs = p.ntiles * cdiv(j->num_freqs, p.FB);
    p.lds = p.ws ? mfma_ws_lds(h16, 16 * p.CT, p.Mp) : mfma_lds(h16, 16 * p.CT, p.Mp);
    return ABF_OK;
}

size_t plain_ut_bytes(int M, int F) { return (size_t)F * round_up(M, 16) * round_up(M, 16) * sizeof(float2); }

// sizes and enums only (enough for workspace queries)
int validate_job(const abf_grid_job *j)
{
    if (!j) return fail(ABF_ERR_INVALID, "null job");
    if (j->geom.num_mics < 1 || j->geom.num_grid < 0 || j->num_freqs < 0)
        return fail(ABF_ERR_INVALID, "bad job sizes");
    if (int r = check_steer(j->steer_type)) return r;
    if (int r = check_phase(j->phase_mode)) return r;
    if (j->form < ABF_FORM_PLAIN_F32 || j->form > ABF_FORM_MFMA_F16X2) return fail(ABF_ERR_INVALID, "unknown form");
    if (j->tile_kernel < ABF_MFMA_KERNEL_AUTO || j->tile_kernel > ABF_MFMA_KERNEL_STREAMED_PERSISTENT)
        return fail(ABF_ERR_INVALID, "unknown MFMA kernel variant");
    if (j->tile_cols < 0 || j->tile_freqs < 0) return fail(ABF_ERR_INVALID, "negative tile size");
    return ABF_OK;
}

int validate_job_pointers(const abf_grid_job *j)
{
    if (j->geom.num_grid > 0 && j->num_freqs > 0) {
        if (!j->csm || !j->wavenumber || !j->geom.gpos || !j->geom.mpos || !j->out)
            return fail(ABF_ERR_INVALID, "null device pointer in job");
    }
    return ABF_OK;
}

size_t job_workspace(const abf_grid_job *j)
{
    if (j->form == ABF_FORM_PLAIN_F32) return align256(plain_ut_bytes(j->geom.num_mics, j->num_freqs));
    MfPlan p;
    if (plan_mfma(j, p)) return 0;
    return align256(p.prep_bytes);
}

// descriptors travel to the device as kernel arguments of this copy kernel, so a launch
// sequence needs no host-to-device memcpy (and can be captured into a graph)
constexpr int kPack = 16;
struct JobPack {
    MfJob jobs[kPack];
    int n;
};

// job descriptors from kernel arguments: no host copy, no memset node, so the sequence can be
// captured in a graph
__global__ void write_jobs_kernel(JobPack pack, MfJob *dst)
{
    const int i = threadIdx.x;
    if (i < pack.n) dst[i] = pack.jobs[i];
}

// batch workspace head: the job descriptors
size_t desc_bytes(int num_jobs) { return align256((size_t)std::max(num_jobs, 1) * sizeof(MfJob)); }

__global__ void set_double_kernel(double *d, double v)
{
    if (threadIdx.x == 0) d[0] = v;
}

__global__ void zero_c64_kernel(float2 *p)
{
    if (threadIdx.x == 0) p[0] = make_float2(0.f, 0.f);
}

__global__ void twiddle_kernel(float2 *tw, int bs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= bs / 2) return;
    double s, c;
    sincospi(2.0 * j / bs, &s, &c);
    tw[j] = make_float2((float)c, (float)-s);
}

template <int B, int CW, int V, int TT, bool PF>
void launch_mfma(const MfJob *dev, int n, int64_t wgs, size_t lds, hipStream_t s)
{
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_mfma_kernel<B, CW, V, TT, PF>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax);
    });
    hipLaunchKernelGGL((bf_mfma_kernel<B, CW, V, TT, PF>), dim3((unsigned)wgs), dim3(512), lds, s, dev, n);
}

template <int B, int V> void launch_mfma_ws(const MfJob *dev, int n, int64_t wgs, size_t lds, hipStream_t s)
{
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_mfma_ws_kernel<B, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLdsMax);
    });
    hipLaunchKernelGGL((bf_mfma_ws_kernel<B, V>), dim3((unsigned)wgs), dim3(512), lds, s, dev, n);
}

template <int B, int V>
void launch_mfma_form(int CW, int TT, bool pf, bool ws, const MfJob *dev, int n, int64_t wgs, size_t lds,
                      hipStream_t s)
{
    if constexpr (B != kPrF16X2) {
        if (ws) launch_mfma_ws<B, V>(dev, n, wgs, lds, s);
    }
    if constexpr (B == kPrBF16X3) {
        if (pf) {
            if (TT == 2) launch_mfma<B, 1, V, 2, true>(dev, n, wgs, lds, s);
            else if (CW == 1) launch_mfma<B, 1, V, 1, true>(dev, n, wgs, lds, s);
            else launch_mfma<B, 2, V, 1, true>(dev, n, wgs, lds, s);
            return;
        }
    }
    if (ws) return;
    if (TT == 2 && CW == 1) launch_mfma<B, 1, V, 2, false>(dev, n, wgs, lds, s);
    else if (TT == 2) launch_mfma<B, 2, V, 2, false>(dev, n, wgs, lds, s);
    else if (CW == 1) launch_mfma<B, 1, V, 1, false>(dev, n, wgs, lds, s);
    else if (CW == 2) launch_mfma<B, 2, V, 1, false>(dev, n, wgs, lds, s);
    else launch_mfma<B, 4, V, 1, false>(dev, n, wgs, lds, s);
}

// steering-generation variant of a job: fast trig (bit 0) | weight kind << 1
int mfma_variant(const abf_grid_job *j)
{
    return ((j->phase_mode & kFastTrigFlag) ? 1 : 0) | (weight_kind(j->steer_type) << 1);
}

template <int B>
void launch_mfma_var(int var, int CW, int TT, bool pf, bool ws, const MfJob *dev, int n, int64_t wgs, size_t lds,
                     hipStream_t s)
{
    switch (var) {
    case 0: launch_mfma_form<B, 0>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    case 1: launch_mfma_form<B, 1>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    case 2: launch_mfma_form<B, 2>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    case 3: launch_mfma_form<B, 3>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    case 4: launch_mfma_form<B, 4>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    default: launch_mfma_form<B, 5>(CW, TT, pf, ws, dev, n, wgs, lds, s); break;
    }
}

// compute units of the current device (looked up once per device)
int cu_count()
{
    static std::atomic<int> cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// PS = false: one work item per workgroup; true: a persistent grid of one workgroup per CU (the
// kernel's LDS allows one) walking the wgs items
template <int NK, int WK, bool PS> void launch_stream_ps(const MfJob *dev, int n, int64_t wgs, size_t lds, hipStream_t s)
{
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_mfma_stream_kernel<NK, WK, PS>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax);
    });
    const int64_t grid = PS ? std::min<int64_t>(wgs, cu_count()) : wgs;
    hipLaunchKernelGGL((bf_mfma_stream_kernel<NK, WK, PS>), dim3((unsigned)grid), dim3(kStWaves * 64), lds, s, dev,
                       n, wgs);
}

template <int NK, int WK> void launch_stream(const MfJob *dev, int n, int64_t wgs, size_t lds, bool ps, hipStream_t s)
{
    if (ps) launch_stream_ps<NK, WK, true>(dev, n, wgs, lds, s);
    else launch_stream_ps<NK, WK, false>(dev, n, wgs, lds, s);
}

template <int WK>
void launch_stream_nk(int nk, const MfJob *dev, int n, int64_t wgs, size_t lds, bool ps, hipStream_t s)
{
    switch (nk) {
    case 2: launch_stream<2, WK>(dev, n, wgs, lds, ps, s); break;
    case 3: launch_stream<3, WK>(dev, n, wgs, lds, ps, s); break;
    case 4: launch_stream<4, WK>(dev, n, wgs, lds, ps, s); break;
    case 5: launch_stream<5, WK>(dev, n, wgs, lds, ps, s); break;
    case 6: launch_stream<6, WK>(dev, n, wgs, lds, ps, s); break;
    case 7: launch_stream<7, WK>(dev, n, wgs, lds, ps, s); break;
    default: launch_stream<8, WK>(dev, n, wgs, lds, ps, s); break;
    }
}

template <int NK, int WK> void launch_wide(const MfJob *dev, int n, int64_t wgs, size_t lds, hipStream_t s)
{
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_mfma_wide_kernel<NK, WK>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax);
    });
    hipLaunchKernelGGL((bf_mfma_wide_kernel<NK, WK>), dim3((unsigned)wgs), dim3(kWdWaves * 64), lds, s, dev, n);
}

template <int WK> void launch_wide_nk(int nk, const MfJob *dev, int n, int64_t wgs, size_t lds, hipStream_t s)
{
    switch (nk) {
    case 2: launch_wide<2, WK>(dev, n, wgs, lds, s); break;
    case 3: launch_wide<3, WK>(dev, n, wgs, lds, s); break;
    case 4: launch_wide<4, WK>(dev, n, wgs, lds, s); break;
    case 5: launch_wide<5, WK>(dev, n, wgs, lds, s); break;
    case 6: launch_wide<6, WK>(dev, n, wgs, lds, s); break;
    case 7: launch_wide<7, WK>(dev, n, wgs, lds, s); break;
    case 8: launch_wide<8, WK>(dev, n, wgs, lds, s); break;
    case 9: launch_wide<9, WK>(dev, n, wgs, lds, s); break;
    case 10: launch_wide<10, WK>(dev, n, wgs, lds, s); break;
    case 11: launch_wide<11, WK>(dev, n, wgs, lds, s); break;
    case 12: launch_wide<12, WK>(dev, n, wgs, lds, s); break;
    case 13: launch_wide<13, WK>(dev, n, wgs, lds, s); break;
    case 14: launch_wide<14, WK>(dev, n, wgs, lds, s); break;
    case 15: launch_wide<15, WK>(dev, n, wgs, lds, s); break;
    default: launch_wide<16, WK>(dev, n, wgs, lds, s); break;
    }
}

void launch_mfma_group(int prec, int CW, int TT, bool pf, bool ws, int var, const MfJob *dev, int n, int64_t wgs,
                       size_t lds, hipStream_t s)
{
    if (prec == kPrBF16X3) launch_mfma_var<kPrBF16X3>(var, CW, TT, pf, ws, dev, n, wgs, lds, s);
    else if (prec == kPrF16X2) launch_mfma_var<kPrF16X2>(var, CW, TT, pf, ws, dev, n, wgs, lds, s);
    else launch_mfma_var<kPrF32>(var, CW, TT, pf, ws, dev, n, wgs, lds, s);
}

// The plain kernel with microphone chunks of JC (ABF_PLAIN_JC = 64, 128 or 256): JC x 512
// bytes of LDS per workgroup of 64 grid points (LU: plus 4 x JC x 128 bytes of U blocks); grid
// (ceil(G / 64), F).
template <int JC, bool LU> static void launch_plain_jc(const PlainArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    constexpr size_t kMaxLds = JC * 64 * sizeof(float2) + (LU ? 4 * JC * 16 * sizeof(float2) : 0);
    static std::once_flag attr; // kernel attribute set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_plain_kernel<JC, LU>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kMaxLds);
    });
    const size_t lds = LU ? kMaxLds : (size_t)std::min(JC, a.Mp) * 64 * sizeof(float2);
    hipLaunchKernelGGL((bf_plain_kernel<JC, LU>), dim3(gx, gy), dim3(256), lds, s, a);
}

// ABF_PLAIN_LDSU=1 stages U in LDS (measured slower than the scalar-cache broadcast, kept for
// comparison; 64-microphone chunks only)
static void launch_plain(const PlainArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    const int jc = env_int("ABF_PLAIN_JC", kPlainDefaultJC);
    if (env_int("ABF_PLAIN_LDSU", 0)) launch_plain_jc<64, true>(a, gx, gy, s);
    else if (jc <= 64) launch_plain_jc<64, false>(a, gx, gy, s);
    else if (jc <= 128) launch_plain_jc<128, false>(a, gx, gy, s);
    else launch_plain_jc<256, false>(a, gx, gy, s);
}

void prep_plain(const abf_grid_job *j, void *ws, hipStream_t s)
{
    const int M = j->geom.num_mics, Mp = round_up(M, 16), F = j->num_freqs;
    hipLaunchKernelGGL(prep_plain_kernel, dim3((unsigned)cdiv((int64_t)Mp * Mp, 256), F), dim3(256), 0, s,
                       reinterpret_cast<const float2 *>(j->csm), M, Mp, j->r_diag, reinterpret_cast<float2 *>(ws));
}

int run_plain_only(const abf_grid_job *j, void *ws, hipStream_t s)
{
    const int M = j->geom.num_mics, Mp = round_up(M, 16), F = j->num_freqs;
    const int64_t G = j->geom.num_grid;
    float2 *ut = reinterpret_cast<float2 *>(ws);
    PlainArgs a{};
    a.ut = ut;
    a.gpos = j->geom.gpos;
    a.mpos = j->geom.mpos;
    a.wavenumber = j->wavenumber;
    a.out = j->out;
    a.out_norm = j->out_norm;
    for (int i = 0; i < 3; ++i) a.ref[i] = j->geom.ref[i];
    a.ref_dist = j->geom.ref_dist;
    a.norm = j->norm_factor;
    a.G = G;
    a.M = M;
    a.Mp = Mp;
    a.steer_type = j->steer_type;
    a.phase_mode = j->phase_mode & kPhaseMask;
    a.fast_trig = (j->phase_mode & kFastTrigFlag) != 0;
    a.clip = j->clip_negative;
    launch_plain(a, (unsigned)cdiv(G, 64), (unsigned)F, s);
    return hip_status("plain beamformer launch");
}

int run_plain(const abf_grid_job *j, void *ws, hipStream_t s)
{
    prep_plain(j, ws, s);
    return run_plain_only(j, ws, s);
}

MfJob make_mfjob(const abf_grid_job *j, const MfPlan &p, const void *prep, int64_t wg_begin)
{
    MfJob d{};
    d.prep = prep;
    d.uscale = p.prec == kPrF16X2 ? reinterpret_cast<const double *>(static_cast<const char *>(prep) + p.uscale_off)
                                  : nullptr;
    d.gpos = j->geom.gpos;
    d.mpos = j->geom.mpos;
    d.wavenumber = j->wavenumber;
    d.out = j->out;
    d.out_norm = j->out_norm;
    for (int i = 0; i < 3; ++i) d.ref[i] = j->geom.ref[i];
    d.ref_dist = j->geom.ref_dist;
    d.norm = j->norm_factor;
    d.G = j->geom.num_grid;
    d.wg_begin = wg_begin;
    d.M = j->geom.num_mics;
    d.Mp = p.Mp;
    d.CT = p.CT;
    d.F = j->num_freqs;
    d.FB = p.FB;
    d.ntiles = (int)p.ntiles;
    d.steer_type = j->steer_type;
    d.phase_mode = j->phase_mode & kPhaseMask;
    d.fast_trig = (j->phase_mode & kFastTrigFlag) != 0;
    d.clip = j->clip_negative;
    return d;
}

void launch_prep_mfma(const abf_grid_job *j, const MfPlan &p, void *prep, hipStream_t s)
{
    double *uscale = nullptr;
    if (p.prec == kPrF16X2) {
        uscale = reinterpret_cast<double *>(static_cast<char *>(prep) + p.uscale_off);
        hipLaunchKernelGGL(uscale_kernel, dim3(j->num_freqs), dim3(256), 0, s,
                           reinterpret_cast<const float2 *>(j->csm), j->geom.num_mics, j->r_diag, uscale);
    }
    hipLaunchKernelGGL(prep_mfma_kernel, dim3(p.nUp, j->num_freqs), dim3(128), 0, s,
                       reinterpret_cast<const float2 *>(j->csm), j->geom.num_mics, p.Mp, j->r_diag, p.prec, uscale,
                       reinterpret_cast<char *>(prep));
}

} // namespace

int abf::set_error(int code, const char *msg) { return fail(code, msg); }

extern "C" {

const char *abf_last_error(void) { return g_err.c_str(); }
const char *abf_version(void) { return "abf 0.1 (gfx950)"; }

size_t abf_beamform_workspace_size(const abf_grid_job *job)
{
    if (validate_job(job)) return 0;
    return abf_beamform_batch_workspace_size(job, 1, nullptr);
}

int abf_beamform_plan(const abf_grid_job *job, int32_t *tile_kernel, int32_t *tile_cols, int32_t *tile_freqs,
                      int64_t *workgroups)
{
    if (int r = validate_job(job)) return r;
    int32_t k = 0, c = 64, f = 1;
    int64_t w = 0;
    if (job->form == ABF_FORM_PLAIN_F32) {
        w = cdiv(job->geom.num_grid, 64) * std::max(job->num_freqs, 0);
    } else {
        MfPlan p;
        if (int r = plan_mfma(job, p)) return r;
        if (p.st)
            k = p.wide ? ABF_MFMA_KERNEL_STREAMED_WIDE
                       : (p.ps ? ABF_MFMA_KERNEL_STREAMED_PERSISTENT : ABF_MFMA_KERNEL_STREAMED);
        else if (p.ws) k = ABF_MFMA_KERNEL_SPECIALISED;
        else if (p.pf) k = p.TT == 2 ? ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS : ABF_MFMA_KERNEL_PIPELINED;
        else k = p.TT == 2 ? ABF_MFMA_KERNEL_ROW_PAIRS : ABF_MFMA_KERNEL_ALTERNATING;
        c = 16 * p.CT;
        f = p.FB;
        w = p.wgs;
    }
    if (tile_kernel) *tile_kernel = k;
    if (tile_cols) *tile_cols = c;
    if (tile_freqs) *tile_freqs = f;
    if (workgroups) *workgroups = w;
    return ABF_OK;
}

size_t abf_beamform_batch_workspace_size(const abf_grid_job *jobs, int32_t num_jobs, size_t *offsets)
{
    size_t off = desc_bytes(num_jobs);
    for (int i = 0; i < num_jobs; ++i) {
        if (validate_job(&jobs[i])) return 0;
        if (offsets) offsets[i] = off;
        off += job_workspace(&jobs[i]);
    }
    return off;
}

} // extern "C"

namespace {
constexpr int kPrep = 1, kRun = 2;

int beamform_batch_impl(const abf_grid_job *jobs, int32_t num_jobs, void *workspace, size_t workspace_bytes,
                        void *stream, int mode)
{
    if (num_jobs < 0 || (num_jobs > 0 && !jobs)) return fail(ABF_ERR_INVALID, "bad job list");
    const hipStream_t s = S(stream);
    std::vector<size_t> offs(num_jobs);
    const size_t need = abf_beamform_batch_workspace_size(jobs, num_jobs, offs.data());
    if (num_jobs > 0 && need == 0) return g_err.empty() ? fail(ABF_ERR_INVALID, "invalid job") : ABF_ERR_INVALID;
    if (!workspace || workspace_bytes < need)
        return fail(ABF_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
    char *ws = reinterpret_cast<char *>(workspace);
    MfJob *dev_jobs = reinterpret_cast<MfJob *>(ws);

    // plain jobs run one after the other; MFMA jobs are grouped per (form, kernel, CW, steering
    // variant) into one launch
    struct Group {
        int prec;
        int CW, TT;
        bool pf, ws;
        int var;
        std::vector<int> idx;
        bool st, wide, ps;
        int nk; // streamed kernel: blocks of 32 microphones
    };
    std::vector<Group> groups;
    std::vector<MfPlan> plans(num_jobs);
    for (int i = 0; i < num_jobs; ++i) {
        const abf_grid_job *j = &jobs[i];
        if (int r = validate_job_pointers(j)) return r;
        if (j->geom.num_grid == 0 || j->num_freqs == 0) continue;
        if (j->form == ABF_FORM_PLAIN_F32) {
            if (mode & kPrep) prep_plain(j, ws + offs[i], s);
            if (mode & kRun)
                if (int r = run_plain_only(j, ws + offs[i], s)) return r;
            continue;
        }
        if (int r = plan_mfma(j, plans[i])) return r;
        if (mode & kPrep) launch_prep_mfma(j, plans[i], ws + offs[i], s);
        if (!(mode & kRun)) continue;
        const int nk = plans[i].st ? plans[i].nK : 0;
        auto it = std::find_if(groups.begin(), groups.end(),
                               [&](const Group &g) {
                                   return g.prec == plans[i].prec && g.CW == plans[i].CW && g.TT == plans[i].TT &&
                                          g.pf == plans[i].pf && g.ws == plans[i].ws && g.st == plans[i].st &&
                                          g.wide == plans[i].wide && g.ps == plans[i].ps && g.nk == nk &&
                                          g.var == mfma_variant(&jobs[i]);
                               });
        if (it == groups.end()) {
            groups.push_back({plans[i].prec, plans[i].CW, plans[i].TT, plans[i].pf, plans[i].ws,
                              mfma_variant(&jobs[i]), {}, plans[i].st, plans[i].wide, plans[i].ps, nk});
            it = groups.end() - 1;
        }
        it->idx.push_back(i);
    }
    if (int r = hip_status("MFMA re-layout launch")) return r;
    int desc_base = 0;
    for (const Group &gr : groups) {
        int64_t wg = 0;
        size_t lds = 0;
        std::vector<MfJob> descs;
        for (int i : gr.idx) {
            descs.push_back(make_mfjob(&jobs[i], plans[i], ws + offs[i], wg));
            wg += plans[i].wgs;
            lds = std::max(lds, plans[i].lds);
        }
        for (size_t b = 0; b < descs.size(); b += kPack) {
            JobPack pk{};
            pk.n = (int)std::min<size_t>(kPack, descs.size() - b);
            for (int q = 0; q < pk.n; ++q) pk.jobs[q] = descs[b + q];
            hipLaunchKernelGGL(write_jobs_kernel, dim3(1), dim3(64), 0, s, pk, dev_jobs + desc_base + b);
        }
        if (gr.st) {
            const int wk = gr.var >> 1;
            const int nd = (int)descs.size();
            if (gr.wide) {
                if (wk == 0) launch_wide_nk<0>(gr.nk, dev_jobs + desc_base, nd, wg, lds, s);
                else if (wk == 1) launch_wide_nk<1>(gr.nk, dev_jobs + desc_base, nd, wg, lds, s);
                else launch_wide_nk<2>(gr.nk, dev_jobs + desc_base, nd, wg, lds, s);
            } else {
                if (wk == 0) launch_stream_nk<0>(gr.nk, dev_jobs + desc_base, nd, wg, lds, gr.ps, s);
                else if (wk == 1) launch_stream_nk<1>(gr.nk, dev_jobs + desc_base, nd, wg, lds, gr.ps, s);
                else launch_stream_nk<2>(gr.nk, dev_jobs + desc_base, nd, wg, lds, gr.ps, s);
            }
        } else {
            launch_mfma_group(gr.prec, gr.CW, gr.TT, gr.pf, gr.ws, gr.var, dev_jobs + desc_base, (int)descs.size(), wg,
                              lds, s);
        }
        if (int r = hip_status("MFMA beamformer launch")) return r;
        desc_base += (int)descs.size();
    }
    return ABF_OK;
}

int beamform_single(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream, int mode)
{
    if (int r = validate_job(job)) return r;
    if (int r = validate_job_pointers(job)) return r;
    if (job->geom.num_grid == 0 || job->num_freqs == 0) return ABF_OK;
    return beamform_batch_impl(job, 1, workspace, workspace_bytes, stream, mode);
}
} // namespace

extern "C" {

int abf_beamform_batch(const abf_grid_job *jobs, int32_t num_jobs, void *workspace, size_t workspace_bytes,
                       void *stream)
{
    return beamform_batch_impl(jobs, num_jobs, workspace, workspace_bytes, stream, kPrep | kRun);
}

int abf_beamform(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream)
{
    return beamform_single(job, workspace, workspace_bytes, stream, kPrep | kRun);
}

int abf_beamform_prepare(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream)
{
    return beamform_single(job, workspace, workspace_bytes, stream, kPrep);
}

int abf_beamform_run(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream)
{
    return beamform_single(job, workspace, workspace_bytes, stream, kRun);
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// abf_beamform_noalloc: abf_beamform with its workspace in a static device array
// ---------------------------------------------------------------------------------------------
namespace {
// Workspace of abf_beamform_noalloc (job descriptors + re-laid-out CSM of a chunk of bins).  Each
// GPU holds its own copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device (scratch_ptr), never by this name.
static __device__ __align__(256) char scratch[ABF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// device address of the current device's scratch array (looked up once per device)
int scratch_ptr(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(ABF_ERR_HIP, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return fail(ABF_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ABF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return ABF_OK;
}
} // namespace

extern "C" {

int abf_beamform_noalloc(const abf_grid_job *job, void *stream)
{
    if (int r = validate_job(job)) return r;
    if (int r = validate_job_pointers(job)) return r;
    if (job->geom.num_grid == 0 || job->num_freqs == 0) return ABF_OK;
    // bins per chunk: descriptors + the re-laid-out CSM of the chunk must fit the scratch array
    abf_grid_job one = *job;
    one.num_freqs = 1;
    const size_t desc = desc_bytes(1);
    const size_t per_bin = job_workspace(&one);
    if (per_bin == 0) return g_err.empty() ? fail(ABF_ERR_INVALID, "invalid job") : ABF_ERR_INVALID;
    if (desc + per_bin > (size_t)ABF_NOALLOC_SCRATCH_BYTES)
        return fail(ABF_ERR_WORKSPACE, "one bin of this job needs " + std::to_string(desc + per_bin) +
                                           " bytes of scratch; abf_beamform_noalloc has " +
                                           std::to_string((size_t)ABF_NOALLOC_SCRATCH_BYTES));
    // f16x2 keeps its per-frequency scales behind the re-laid-out CSM: one more 256-byte unit
    const size_t room = (size_t)ABF_NOALLOC_SCRATCH_BYTES - desc - 256;
    const int chunk = (int)std::min<size_t>((size_t)job->num_freqs, std::max<size_t>(1, room / per_bin));
    void *ws = nullptr;
    if (int r = scratch_ptr(&ws)) return r;
    const int64_t M = job->geom.num_mics, G = job->geom.num_grid;
    for (int b0 = 0; b0 < job->num_freqs; b0 += chunk) {
        abf_grid_job sub = *job;
        sub.num_freqs = std::min(chunk, job->num_freqs - b0);
        sub.csm = static_cast<const char *>(job->csm) + (size_t)b0 * M * M * 8;
        sub.wavenumber = job->wavenumber + b0;
        sub.out = job->out + (size_t)b0 * G;
        sub.out_norm = job->out_norm ? job->out_norm + (size_t)b0 * G : nullptr;
        if (int r = beamform_batch_impl(&sub, 1, ws, (size_t)ABF_NOALLOC_SCRATCH_BYTES, stream, kPrep | kRun))
            return r;
    }
    return ABF_OK;
}

size_t abf_beamform_table_workspace_size(int32_t num_mics)
{
    return align256(plain_ut_bytes(num_mics, 1)) + 256;
}

int abf_beamform_table(const void *csm, const double *r0, const double *rm, int64_t num_grid, int32_t num_mics,
                       double wavenumber, int32_t steer_type, int32_t r_diag, double norm_factor, int32_t phase_mode,
                       float *out, float *out_norm, void *workspace, size_t workspace_bytes, void *stream)
{
    if (int r = check_steer(steer_type)) return r;
    if (int r = check_phase(phase_mode)) return r;
    if (num_mics < 1 || num_grid < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_grid == 0) return ABF_OK;
    if (!csm || !r0 || !rm || !out) return fail(ABF_ERR_INVALID, "null device pointer");
    const size_t need = abf_beamform_table_workspace_size(num_mics);
    if (!workspace || workspace_bytes < need) return fail(ABF_ERR_WORKSPACE, "workspace too small");
    const hipStream_t s = S(stream);
    const int M = num_mics, Mp = round_up(M, 16);
    char *ws = reinterpret_cast<char *>(workspace);
    float2 *ut = reinterpret_cast<float2 *>(ws);
    double *kdev = reinterpret_cast<double *>(ws + align256(plain_ut_bytes(M, 1)));
    hipLaunchKernelGGL(prep_plain_kernel, dim3((unsigned)cdiv((int64_t)Mp * Mp, 256), 1), dim3(256), 0, s,
                       reinterpret_cast<const float2 *>(csm), M, Mp, r_diag, ut);
    // the wavenumber reaches the device through a one-element fill (no host memcpy)
    hipLaunchKernelGGL(set_double_kernel, dim3(1), dim3(64), 0, s, kdev, wavenumber);
    PlainArgs a{};
    a.ut = ut;
    a.rm_table = rm;
    a.r0_table = r0;
    a.wavenumber = kdev;
    a.out = out;
    a.out_norm = out_norm;
    a.norm = norm_factor;
    a.G = num_grid;
    a.M = M;
    a.Mp = Mp;
    a.steer_type = steer_type;
    a.phase_mode = phase_mode & kPhaseMask;
    a.fast_trig = (phase_mode & kFastTrigFlag) != 0;
    a.clip = 0;
    launch_plain(a, (unsigned)cdiv(num_grid, 64), 1u, s);
    return hip_status("table beamformer launch");
}

size_t abf_beamform_eig_workspace_size(int32_t, int32_t) { return 0; }

int abf_beamform_eig(const double *eigval, const void *eigvec, int32_t num_eig, const double *r0, const double *rm,
                     int64_t num_grid, int32_t num_mics, double wavenumber, int32_t steer_type, int32_t r_diag,
                     double norm_factor, int32_t phase_mode, float *out, float *out_norm, void *, size_t,
                     void *stream)
{
    if (int r = check_steer(steer_type)) return r;
    if (int r = check_phase(phase_mode)) return r;
    if (num_mics < 1 || num_grid < 0 || num_eig < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_grid == 0) return ABF_OK;
    if (!r0 || !rm || !out || (num_eig > 0 && (!eigval || !eigvec))) return fail(ABF_ERR_INVALID, "null pointer");
    EigArgs a{};
    a.eigval = eigval;
    a.eigvec_t = reinterpret_cast<const float2 *>(eigvec);
    a.steer = nullptr;
    a.r0_table = r0;
    a.rm_table = rm;
    a.out = out;
    a.out_norm = out_norm;
    a.wavenumber = wavenumber;
    a.norm = norm_factor;
    a.G = num_grid;
    a.M = num_mics;
    a.K = num_eig;
    a.steer_type = steer_type;
    a.r_diag = r_diag;
    a.phase_mode = phase_mode & kPhaseMask;
    a.fast_trig = (phase_mode & kFastTrigFlag) != 0;
    hipLaunchKernelGGL(bf_eig_kernel, dim3((unsigned)cdiv(num_grid, 256)), dim3(256), 0, S(stream), a);
    return hip_status("eig beamformer launch");
}

int abf_beamform_eig_custom(const double *eigval, const void *eigvec, int32_t num_eig, const void *steer,
                            int64_t num_grid, int32_t num_mics, int32_t r_diag, double norm_factor, float *out,
                            float *out_norm, void *, size_t, void *stream)
{
    if (num_mics < 1 || num_grid < 0 || num_eig < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_grid == 0) return ABF_OK;
    if (!steer || !out || (num_eig > 0 && (!eigval || !eigvec))) return fail(ABF_ERR_INVALID, "null pointer");
    EigArgs a{};
    a.eigval = eigval;
    a.eigvec_t = reinterpret_cast<const float2 *>(eigvec);
    a.steer = reinterpret_cast<const float2 *>(steer);
    a.out = out;
    a.out_norm = out_norm;
    a.norm = norm_factor;
    a.G = num_grid;
    a.M = num_mics;
    a.K = num_eig;
    a.steer_type = ABF_STEER_CLASSIC;
    a.r_diag = r_diag;
    hipLaunchKernelGGL(bf_eig_kernel, dim3((unsigned)cdiv(num_grid, 256)), dim3(256), 0, S(stream), a);
    return hip_status("eig custom beamformer launch");
}

size_t abf_beamform_f64_workspace_size(int32_t num_mics, int32_t num_freqs)
{
    if (num_mics < 1 || num_freqs < 0) return 0;
    return align256((size_t)num_freqs * round_up(num_mics, 16) * round_up(num_mics, 16) * sizeof(double2));
}

int abf_beamform_f64(const abf_geometry *geom, const void *csm, const double *wavenumber, int32_t num_freqs,
                     int32_t steer_type, int32_t r_diag, double norm_factor, int32_t phase_mode, int32_t clip_negative,
                     double *out, double *out_norm, void *workspace, size_t workspace_bytes, void *stream)
{
    if (!geom) return fail(ABF_ERR_INVALID, "null geometry");
    if (int r = check_steer(steer_type)) return r;
    if (int r = check_phase(phase_mode)) return r;
    const int M = geom->num_mics;
    const int64_t G = geom->num_grid;
    if (M < 1 || G < 0 || num_freqs < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_freqs > 65535) return fail(ABF_ERR_INVALID, "at most 65535 frequencies per call");
    if (G == 0 || num_freqs == 0) return ABF_OK;
    if (!geom->gpos || !geom->mpos || !csm || !wavenumber || !out) return fail(ABF_ERR_INVALID, "null pointer");
    if (!workspace || workspace_bytes < abf_beamform_f64_workspace_size(M, num_freqs))
        return fail(ABF_ERR_WORKSPACE, "workspace too small");
    const int Mp = round_up(M, 16);
    hipStream_t s = S(stream);
    double2 *ut = reinterpret_cast<double2 *>(workspace);
    hipLaunchKernelGGL(prep_plain64_kernel, dim3((unsigned)cdiv((int64_t)Mp * Mp, 256), num_freqs), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(csm), M, Mp, r_diag, ut);
    Plain64Args a{};
    a.ut = ut;
    a.gpos = geom->gpos;
    a.mpos = geom->mpos;
    a.wavenumber = wavenumber;
    a.out = out;
    a.out_norm = out_norm;
    for (int i = 0; i < 3; ++i) a.ref[i] = geom->ref[i];
    a.ref_dist = geom->ref_dist;
    a.norm = norm_factor;
    a.G = G;
    a.M = M;
    a.Mp = Mp;
    a.steer_type = steer_type;
    a.phase_mode = phase_mode & kPhaseMask;
    a.clip = clip_negative;
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_plain64_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 131072);
    });
    const size_t lds = (size_t)std::min(128, Mp) * 64 * sizeof(double2);
    hipLaunchKernelGGL(bf_plain64_kernel, dim3((unsigned)cdiv(G, 64), num_freqs), dim3(256), lds, s, a);
    return hip_status("float64 beamformer launch");
}

int abf_beamform_eig_grid(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, const double *eigval,
                          const void *eigvec, int32_t eig_first, int32_t num_eig, int32_t steer_type, int32_t r_diag,
                          double norm_factor, int32_t phase_mode, int32_t clip_negative, float *out, float *out_norm,
                          void *stream)
{
    if (!geom) return fail(ABF_ERR_INVALID, "null geometry");
    if (int r = check_steer(steer_type)) return r;
    if (int r = check_phase(phase_mode)) return r;
    const int M = geom->num_mics;
    const int64_t G = geom->num_grid;
    if (M < 1 || G < 0 || num_freqs < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (eig_first < 0 || num_eig < 0 || (int64_t)eig_first + num_eig > M)
        return fail(ABF_ERR_INVALID, "eigenvalue range outside [0, num_mics)");
    if (num_freqs > 65535) return fail(ABF_ERR_INVALID, "at most 65535 frequencies per call");
    if (G == 0 || num_freqs == 0) return ABF_OK;
    if (!geom->gpos || !geom->mpos || !wavenumber || !out || !eigval || !eigvec)
        return fail(ABF_ERR_INVALID, "null pointer");
    // many eigenvectors without r_diag: the fp32 MFMA kernel, 16-point column tiles, the
    // steering tile [Mp][P] within 144 KiB
    const int Mp4 = (M + 3) & ~3;
    const bool mfma = !r_diag && num_eig >= 16 && (size_t)Mp4 * 16 * sizeof(float2) <= 147456;
    // grid points per workgroup: the steering tile [M][P] stays within 64 KiB where it can
    // (two or more workgroups per CU), within 144 KiB otherwise
    int P = 64;
    if (mfma) {
        // two workgroups per CU where the tile allows (64 KiB), one otherwise
        while (P > 16 && (size_t)Mp4 * P * sizeof(float2) > 65536) P /= 2;
    } else {
        while (P > 4 && (size_t)M * P * sizeof(float2) > 65536) P /= 2;
    }
    const size_t lds = (size_t)(mfma ? Mp4 : M) * P * sizeof(float2);
    if (lds > 147456) return fail(ABF_ERR_UNSUPPORTED, "too many microphones for the eigenvalue form");
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)bf_eig_grid_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 147456);
        (void)hipFuncSetAttribute((const void *)bf_eig_mfma_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 147456);
    });
    EigGridArgs a{};
    a.gpos = geom->gpos;
    a.mpos = geom->mpos;
    a.wavenumber = wavenumber;
    a.eigval = eigval;
    a.eigvec = reinterpret_cast<const float2 *>(eigvec);
    a.out = out;
    a.out_norm = out_norm;
    for (int i = 0; i < 3; ++i) a.ref[i] = geom->ref[i];
    a.ref_dist = geom->ref_dist;
    a.norm = norm_factor;
    a.G = G;
    a.M = M;
    a.e0 = eig_first;
    a.K = num_eig;
    a.P = P;
    a.steer_type = steer_type;
    a.r_diag = r_diag;
    a.phase_mode = phase_mode & kPhaseMask;
    a.fast_trig = (phase_mode & kFastTrigFlag) != 0;
    a.clip = clip_negative;
    if (mfma)
        hipLaunchKernelGGL(bf_eig_mfma_kernel, dim3((unsigned)cdiv(G, P), (unsigned)num_freqs), dim3(256), lds,
                           S(stream), a);
    else
        hipLaunchKernelGGL(bf_eig_grid_kernel, dim3((unsigned)cdiv(G, P), (unsigned)num_freqs), dim3(256), lds,
                           S(stream), a);
    return hip_status("eigenvalue grid beamformer launch");
}

int abf_beamform_custom(const void *csm, const void *steer, int64_t num_grid, int32_t num_mics, int32_t r_diag,
                        double norm_factor, float *out, float *out_norm, void *, size_t, void *stream)
{
    if (num_mics < 1 || num_grid < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_grid == 0) return ABF_OK;
    if (!csm || !steer || !out) return fail(ABF_ERR_INVALID, "null pointer");
    CustomArgs a{};
    a.csm = reinterpret_cast<const float2 *>(csm);
    a.steer = reinterpret_cast<const float2 *>(steer);
    a.out = out;
    a.out_norm = out_norm;
    a.norm = norm_factor;
    a.G = num_grid;
    a.M = num_mics;
    a.r_diag = r_diag;
    hipLaunchKernelGGL(bf_custom_kernel, dim3((unsigned)cdiv(num_grid, 256)), dim3(256), 0, S(stream), a);
    return hip_status("custom beamformer launch");
}

int abf_dist_mat(const double *gpos, const double *mpos, int64_t num_grid, int32_t num_mics, double *rm,
                 void *stream)
{
    if (num_mics < 1 || num_grid < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_grid == 0) return ABF_OK;
    if (!gpos || !mpos || !rm) return fail(ABF_ERR_INVALID, "null pointer");
    const int64_t nmb = cdiv((int64_t)num_mics, 512);
    // points per launch: at most 2^22 workgroups of kDistPts points
    const int64_t per = std::max<int64_t>(1, (int64_t(1) << 22) / nmb) * kDistPts;
    // two distances per 16-byte store when every row starts 16-byte aligned
    const int vec2 = !(num_mics & 1) && aligned16(rm);
    for (int64_t g0 = 0; g0 < num_grid; g0 += per)
        hipLaunchKernelGGL(dist_mat_kernel, dim3((unsigned)(cdiv(std::min(per, num_grid - g0), kDistPts) * nmb)),
                           dim3(256), 0, S(stream), gpos, mpos, num_grid, (int)num_mics, g0, vec2, rm);
    return hip_status("dist_mat launch");
}

int abf_transfer(const double *r0, const double *rm, int64_t num_grid, int32_t num_mics, double wavenumber,
                 int32_t phase_mode, int32_t steer_mode, int32_t out_c128, void *out, void *stream)
{
    if (num_mics < 1 || num_grid < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (steer_mode < 0 || steer_mode > 4) return fail(ABF_ERR_INVALID, "unknown steer mode");
    if (int r = check_phase(phase_mode)) return r;
    if (num_grid == 0) return ABF_OK;
    if (!r0 || !rm || !out) return fail(ABF_ERR_INVALID, "null pointer");
    for (int64_t g0 = 0; g0 < num_grid; g0 += int64_t(1) << 24) { // 2^22 workgroups of 4 points per launch
        const int64_t ng = std::min<int64_t>(int64_t(1) << 24, num_grid - g0);
        const double *r0p = r0 + g0, *rmp = rm + g0 * num_mics;
        void *outp = static_cast<char *>(out) + g0 * num_mics * (out_c128 ? 16 : 8);
        hipLaunchKernelGGL(transfer_kernel, dim3((unsigned)cdiv(ng, 4)), dim3(256), 0, S(stream), r0p, rmp, ng,
                           (int)num_mics, wavenumber, phase_mode & kPhaseMask, steer_mode, out_c128, outp);
    }
    return hip_status("transfer launch");
}

int abf_calc_csm(void *csm, const void *spec, int32_t num_freqs, int32_t num_mics, int32_t is_c128, void *stream)
{
    if (num_mics < 1 || num_freqs < 0) return fail(ABF_ERR_INVALID, "bad sizes");
    if (num_freqs == 0) return ABF_OK;
    if (!csm || !spec) return fail(ABF_ERR_INVALID, "null pointer");
    if (num_mics > 65535) return fail(ABF_ERR_UNSUPPORTED, "more than 65535 microphones");
    const int M = num_mics;
    // complex64: two columns per 16-byte access when every row starts 16-byte aligned
    const int vec2 = !(M & 1) && aligned16(csm) && aligned16(spec);
    for (int f0 = 0; f0 < num_freqs; f0 += 65535) {
        const unsigned nf = (unsigned)std::min(65535, num_freqs - f0);
        if (is_c128)
            hipLaunchKernelGGL(calc_csm_c128_kernel, dim3((unsigned)cdiv(M, 256), (unsigned)M, nf), dim3(256), 0,
                               S(stream), reinterpret_cast<double2 *>(csm), reinterpret_cast<const double2 *>(spec), f0,
                               M);
        else
            hipLaunchKernelGGL(calc_csm_c64_kernel, dim3((unsigned)cdiv(M, 512), (unsigned)M, nf), dim3(256), 0,
                               S(stream), reinterpret_cast<float2 *>(csm), reinterpret_cast<const float2 *>(spec), f0,
                               M, vec2);
    }
    return hip_status("calc_csm launch");
}

// Slices per pass of the CSM update: when the tiles of all bins give too few waves to fill the
// GPU (small arrays, e.g. 64 mics x 65 bins: 130 waves walking 799 blocks each), the blocks of a
// pass are split into up to 16 slices of at least 32 blocks, whose partial sums the workspace
// holds (at most 256 MB) and welch_reduce_kernel adds in a fixed order.
extern "C++" { // launch helpers of the block FFT (templates, inside the C-linkage entry points)
// Windowed block FFT launch shape: the radix kernel (block sizes 16 ... 4096; ABF_FFT_RADIX=0
// selects the radix-2 kernel for comparison) with the largest power-of-two pair count P that the
// channels can use, P bs <= 16 x its threads and LDS within kFftRadixLdsMax; the radix-2 kernel
// with at most 64 KB of LDS otherwise.
struct FftShape {
    bool radix;
    int P, logP;
    size_t lds;
    int bpw_max; // radix kernel: at most this many consecutive blocks per workgroup
};

static FftShape fft_shape(int bs, int logbs, int M)
{
    FftShape f{};
    f.radix = logbs >= kFftRadixMinLog && logbs <= kFftRadixMaxLog && env_int("ABF_FFT_RADIX", 1) != 0;
    if (f.radix) {
        int P = 1, lp = 0;
        while (2 * P * bs <= 16 * kFftRadixThreads && P < (M + 1) / 2 && welch_fft_lds_bytes(bs, 2 * P) <= kFftRadixLdsMax) {
            P *= 2;
            ++lp;
        }
        f.P = P;
        f.logP = lp;
        f.lds = welch_fft_lds_bytes(bs, P);
        f.bpw_max = std::max(1, env_int("ABF_FFT_BPW", 8));
    } else {
        // measured 2.36 vs 2.97 ms for 128 KB on 499 blocks x 512 channels of 2048 samples;
        // 16384-sample blocks need 128 KB
        f.P = std::max(1, std::min({env_int("ABF_FFT_P", 16), 8192 / bs, (M + 1) / 2}));
        f.logP = 0;
        f.lds = (size_t)f.P * bs * sizeof(float2);
    }
    return f;
}

template <int LOGN, bool VEC>
static void launch_fft_radix_vec(unsigned nwg, size_t lds, hipStream_t s, const WelchFftArgs &a)
{
    static std::once_flag attr; // kernel attribute set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)welch_fft_radix_kernel<LOGN, VEC>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFftRadixLdsMax);
    });
    hipLaunchKernelGGL((welch_fft_radix_kernel<LOGN, VEC>), dim3(nwg), dim3(kFftRadixThreads), lds, s, a);
}

// the 8-byte-load / 16-byte-store form where M is even and the buffers are aligned for it
template <int LOGN> static void launch_fft_radix_one(unsigned nwg, size_t lds, hipStream_t s, const WelchFftArgs &a)
{
    const bool vec = !(a.M & 1) && !(reinterpret_cast<uintptr_t>(a.time) & 7) &&
                     !(reinterpret_cast<uintptr_t>(a.spec) & 15);
    if (vec) launch_fft_radix_vec<LOGN, true>(nwg, lds, s, a);
    else launch_fft_radix_vec<LOGN, false>(nwg, lds, s, a);
}

// one launch of the block FFT kernel that fft_shape chose over the nbl blocks a.bl0 ... of a pass
// and ng channel groups.  The radix kernel takes up to bpw_max consecutive blocks per workgroup
// while at least 2048 workgroups remain (eight per CU).
static void launch_fft(const FftShape &f, int ng, int nbl, hipStream_t s, WelchFftArgs a)
{
    a.nbl = nbl;
    if (!f.radix) {
        static std::once_flag attr;
        std::call_once(attr, [] {
            (void)hipFuncSetAttribute((const void *)welch_fft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      131072);
        });
        hipLaunchKernelGGL(welch_fft_kernel, dim3((unsigned)((int64_t)ng * nbl)), dim3(512), f.lds, s, a);
        return;
    }
    a.bpw = (int)std::max<int64_t>(1, std::min<int64_t>(f.bpw_max, (int64_t)ng * nbl / 2048));
    const unsigned nwg = (unsigned)((int64_t)ng * cdiv(nbl, a.bpw));
    switch (a.logbs) {
    case 4: launch_fft_radix_one<4>(nwg, f.lds, s, a); break;
    case 5: launch_fft_radix_one<5>(nwg, f.lds, s, a); break;
    case 6: launch_fft_radix_one<6>(nwg, f.lds, s, a); break;
    case 7: launch_fft_radix_one<7>(nwg, f.lds, s, a); break;
    case 8: launch_fft_radix_one<8>(nwg, f.lds, s, a); break;
    case 9: launch_fft_radix_one<9>(nwg, f.lds, s, a); break;
    case 10: launch_fft_radix_one<10>(nwg, f.lds, s, a); break;
    case 11: launch_fft_radix_one<11>(nwg, f.lds, s, a); break;
    default: launch_fft_radix_one<12>(nwg, f.lds, s, a); break;
    }
}
} // extern "C++"

static int welch_split(int M, int bs, int blocks_per_pass)
{
    const int64_t F = bs / 2 + 1;
    const int64_t waves = cdiv((int64_t)M, 64) * cdiv((int64_t)M, 64) * F;
    int sp = 1;
    while (sp < 16 && waves * sp * 2 <= 8192 && blocks_per_pass >= 64 * sp &&
           (int64_t)sp * 2 * F * M * M * 8 <= (int64_t(256) << 20))
        sp *= 2;
    return sp;
}

size_t abf_welch_csm_workspace_size(int32_t num_mics, int32_t block_size, int32_t blocks_per_pass)
{
    if (num_mics < 1 || block_size < 2 || blocks_per_pass < 1) return 0;
    const int sp = welch_split(num_mics, block_size, blocks_per_pass);
    const size_t parts = sp > 1 ? align256((size_t)sp * (block_size / 2 + 1) * num_mics * num_mics * sizeof(float2)) : 0;
    return align256((size_t)block_size / 2 * sizeof(float2)) +
           align256((size_t)(block_size / 2 + 1) * blocks_per_pass * num_mics * sizeof(float2)) + parts;
}

int abf_welch_csm(const float *time, int64_t num_samples, int32_t num_mics, int32_t block_size, int32_t hop,
                  int32_t num_blocks, const float *window, double scale, void *csm_out, int32_t blocks_per_pass,
                  void *workspace, size_t workspace_bytes, void *stream)
{
    const int bs = block_size;
    if (num_mics < 1 || bs < 2 || (bs & (bs - 1)) || hop < 1 || num_blocks < 0 || blocks_per_pass < 1)
        return fail(ABF_ERR_INVALID, "bad Welch parameters");
    if (bs > 16384) return fail(ABF_ERR_UNSUPPORTED, "block_size above 16384 is not supported by the LDS FFT");
    if (num_blocks > 0 && (int64_t)(num_blocks - 1) * hop + bs > num_samples)
        return fail(ABF_ERR_INVALID, "blocks exceed the time data");
    if (!csm_out || (num_blocks > 0 && (!time || !window))) return fail(ABF_ERR_INVALID, "null pointer");
    const size_t need = abf_welch_csm_workspace_size(num_mics, bs, blocks_per_pass);
    if (!workspace || workspace_bytes < need) return fail(ABF_ERR_WORKSPACE, "workspace too small");
    const hipStream_t s = S(stream);
    const int M = num_mics, F = bs / 2 + 1;
    int logbs = 0;
    while ((1 << logbs) < bs) ++logbs;
    char *ws = reinterpret_cast<char *>(workspace);
    float2 *tw = reinterpret_cast<float2 *>(ws);
    float2 *spec = reinterpret_cast<float2 *>(ws + align256((size_t)bs / 2 * sizeof(float2)));
    const int nsplit = welch_split(M, bs, blocks_per_pass);
    float2 *part = reinterpret_cast<float2 *>(
        ws + align256((size_t)bs / 2 * sizeof(float2)) +
        align256((size_t)(bs / 2 + 1) * blocks_per_pass * M * sizeof(float2)));
    if (num_blocks == 0) { // no pass writes the CSM: it is zero
        const int64_t n4 = (int64_t)F * M * M / 2; // complex64 pairs per float4
        const int64_t total = (int64_t)F * M * M;
        hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)cdiv(cdiv(total, 2), 256)), dim3(256), 0, s,
                           reinterpret_cast<float4 *>(csm_out), n4);
        if (total & 1) // odd element count: clear the last complex64 separately
            hipLaunchKernelGGL(zero_c64_kernel, dim3(1), dim3(64), 0, s,
                               reinterpret_cast<float2 *>(csm_out) + (total - 1));
    }
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)cdiv(bs / 2, 256)), dim3(256), 0, s, tw, bs);
    const FftShape fsh = fft_shape(bs, logbs, M);
    const int P = fsh.P;
    const int nT = (M + 63) / 64;
    // 64 x 64 tiles with the next step's loads in flight across the MFMAs (three waves per SIMD)
    // from 256 mics: CSM of the 512-mic, 1025-bin workload 11.8 ms against 13.9 ms for 32 x 64
    // tiles (12.4 ms without that prefetch, five waves per SIMD); 32 x 64 tiles below, where
    // twice the waves matter more than the loads per MFMA
    const int herk_tr = env_int("ABF_HERK_TR", M >= 256 ? 4 : 2) == 4 ? 4 : 2;
    for (int b0 = 0; b0 < num_blocks; b0 += blocks_per_pass) {
        const int nb = std::min(blocks_per_pass, num_blocks - b0);
        WelchFftArgs a{};
        a.time = time;
        a.window = window;
        a.twiddle = tw;
        a.spec = spec;
        a.N = num_samples;
        a.M = M;
        a.bs = bs;
        a.logbs = logbs;
        a.hop = hop;
        a.b0 = b0;
        a.nb = nb;
        a.P = P;
        a.logP = fsh.logP;
        const int ng = (int)cdiv(M, 2 * P);
        const int per_launch = (int)std::max<int64_t>(1, kFftMaxWorkgroups / ng);
        for (int c0 = 0; c0 < nb; c0 += per_launch) {
            a.bl0 = c0;
            launch_fft(fsh, ng, std::min(per_launch, nb - c0), s, a);
        }
        // one-wave workgroups: 2^24 per launch stays far below the 2^32 work-items a dispatch counts
        const int64_t tiles = (int64_t)(herk_tr == 2 ? cdiv(M, 32) : nT) * nT * nsplit;
        const int fper = (int)std::max<int64_t>(1, (int64_t(1) << 24) / tiles);
        for (int fc = 0; fc < F; fc += fper) {
            const unsigned nwg = (unsigned)(tiles * std::min(fper, F - fc));
            if (herk_tr == 2)
                hipLaunchKernelGGL(welch_herk_kernel<2>, dim3(nwg), dim3(64), 0, s, spec, nb, M, nT, fc, (int)(b0 > 0),
                                   reinterpret_cast<float2 *>(csm_out), nsplit, F, part);
            else
                hipLaunchKernelGGL(welch_herk_kernel<4>, dim3(nwg), dim3(64), 0, s, spec, nb, M, nT, fc, (int)(b0 > 0),
                                   reinterpret_cast<float2 *>(csm_out), nsplit, F, part);
        }
        if (nsplit > 1)
            hipLaunchKernelGGL(welch_reduce_kernel, dim3((unsigned)cdiv(M, 256), (unsigned)M, (unsigned)F), dim3(256), 0,
                               s, part, nsplit, F, M, (int)(b0 > 0), reinterpret_cast<float2 *>(csm_out));
    }
    hipLaunchKernelGGL(welch_finalize_kernel, dim3((unsigned)(cdiv(M, 32) * cdiv(M, 32)), F), dim3(256), 0, s,
                       reinterpret_cast<float2 *>(csm_out), M, (float)scale);
    return hip_status("Welch CSM launch");
}

size_t abf_rfft_workspace_size(int32_t block_size)
{
    if (block_size < 2) return 0;
    return align256((size_t)block_size / 2 * sizeof(float2));
}

int abf_rfft_blocks(const float *time, int64_t num_samples, int32_t num_mics, int32_t block_size, int32_t hop,
                    int32_t num_blocks, const float *window, const float *bin_scale, void *out, void *workspace,
                    size_t workspace_bytes, void *stream)
{
    const int bs = block_size;
    if (num_mics < 1 || bs < 2 || (bs & (bs - 1)) || hop < 1 || num_blocks < 0)
        return fail(ABF_ERR_INVALID, "bad FFT block parameters (block_size must be a power of two >= 2)");
    if (bs > 16384) return fail(ABF_ERR_UNSUPPORTED, "block_size above 16384 is not supported by the LDS FFT");
    if (num_blocks > 0 && (int64_t)(num_blocks - 1) * hop + bs > num_samples)
        return fail(ABF_ERR_INVALID, "blocks exceed the time data");
    if (num_blocks == 0) return ABF_OK;
    if (!out || !time || !window) return fail(ABF_ERR_INVALID, "null pointer");
    if (!workspace || workspace_bytes < abf_rfft_workspace_size(bs)) return fail(ABF_ERR_WORKSPACE, "workspace too small");
    const hipStream_t s = S(stream);
    const int M = num_mics;
    int logbs = 0;
    while ((1 << logbs) < bs) ++logbs;
    float2 *tw = reinterpret_cast<float2 *>(workspace);
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)cdiv(bs / 2, 256)), dim3(256), 0, s, tw, bs);
    const FftShape fsh = fft_shape(bs, logbs, M);
    const int P = fsh.P;
    const int ng = (int)cdiv(M, 2 * P);
    const int per_launch = (int)std::min<int64_t>(65535, std::max<int64_t>(1, kFftMaxWorkgroups / ng));
    for (int b0 = 0; b0 < num_blocks; b0 += per_launch) {
        WelchFftArgs a{};
        a.time = time;
        a.window = window;
        a.twiddle = tw;
        a.spec = reinterpret_cast<float2 *>(out);
        a.bin_scale = bin_scale;
        a.N = num_samples;
        a.M = M;
        a.bs = bs;
        a.logbs = logbs;
        a.hop = hop;
        a.b0 = b0;
        a.nb = std::min(per_launch, num_blocks - b0);
        a.P = P;
        a.logP = fsh.logP;
        a.block_major = 1;
        launch_fft(fsh, ng, a.nb, s, a);
    }
    return hip_status("block FFT launch");
}

size_t abf_irfft_workspace_size(int32_t block_size) { return abf_rfft_workspace_size(block_size); }

int abf_irfft_blocks(const void *spec, int32_t num_blocks, int32_t num_mics, int32_t block_size, float *out,
                     void *workspace, size_t workspace_bytes, void *stream)
{
    const int bs = block_size;
    if (num_mics < 1 || bs < 2 || (bs & (bs - 1)) || num_blocks < 0)
        return fail(ABF_ERR_INVALID, "bad inverse FFT parameters (block_size must be a power of two >= 2)");
    if (bs > 16384) return fail(ABF_ERR_UNSUPPORTED, "block_size above 16384 is not supported by the LDS FFT");
    if (num_blocks == 0) return ABF_OK;
    if (!spec || !out) return fail(ABF_ERR_INVALID, "null pointer");
    if (!workspace || workspace_bytes < abf_irfft_workspace_size(bs)) return fail(ABF_ERR_WORKSPACE, "workspace too small");
    const hipStream_t s = S(stream);
    const int M = num_mics;
    int logbs = 0;
    while ((1 << logbs) < bs) ++logbs;
    float2 *tw = reinterpret_cast<float2 *>(workspace);
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)cdiv(bs / 2, 256)), dim3(256), 0, s, tw, bs);
    static std::once_flag attr; // kernel attributes set once, safe under concurrent first calls
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)irfft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 131072);
    });
    const int P = std::max(1, std::min({16, 8192 / bs, (M + 1) / 2}));
    const int ng = (int)cdiv(M, 2 * P);
    const int per_launch = (int)std::min<int64_t>(65535, std::max<int64_t>(1, kFftMaxWorkgroups / ng));
    for (int b0 = 0; b0 < num_blocks; b0 += per_launch) {
        const int nb = std::min(per_launch, num_blocks - b0);
        hipLaunchKernelGGL(irfft_kernel, dim3((unsigned)((int64_t)ng * nb)), dim3(512), (size_t)P * bs * sizeof(float2),
                           s, reinterpret_cast<const float2 *>(spec), M, bs, logbs, b0, P, tw, out);
    }
    return hip_status("inverse block FFT launch");
}

int abf_auto_power(const void *spec, int64_t n, double scale, float *out, void *stream)
{
    if (n < 0) return fail(ABF_ERR_INVALID, "negative size");
    if (n == 0) return ABF_OK;
    if (!spec || !out) return fail(ABF_ERR_INVALID, "null pointer");
    hipLaunchKernelGGL(auto_power_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, S(stream),
                       reinterpret_cast<const float2 *>(spec), n, (float)scale, out);
    return hip_status("auto power launch");
}

int abf_cross_power(const void *spec, int64_t num_snapshots, int32_t num_freqs, int32_t num_mics, double scale,
                    void *out, void *stream)
{
    if (num_snapshots < 0 || num_freqs < 0 || num_mics < 1) return fail(ABF_ERR_INVALID, "bad sizes");
    const int64_t nbf = num_snapshots * num_freqs;
    if (nbf == 0) return ABF_OK;
    if (!spec || !out) return fail(ABF_ERR_INVALID, "null pointer");
    const int64_t mm = (int64_t)num_mics * num_mics;
    const int64_t gz = cdiv(nbf, 65535);
    if (gz > 65535) return fail(ABF_ERR_UNSUPPORTED, "too many snapshots x bins for one call");
    hipLaunchKernelGGL(cross_power_kernel, dim3((unsigned)cdiv(mm, 256), (unsigned)std::min<int64_t>(nbf, 65535),
                                                (unsigned)gz),
                       dim3(256), 0, S(stream), reinterpret_cast<const float2 *>(spec), nbf, num_mics, (float)scale,
                       reinterpret_cast<float2 *>(out));
    return hip_status("cross power launch");
}

} // extern "C"
