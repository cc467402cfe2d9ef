// Batched Hermitian eigendecomposition of a resident CSM, for the eigenvalue beamformers.
//
// The reference decomposes every frequency slice with numpy.linalg.eigh in the CSM's precision
// (PowerSpectra._get__evd, acoular/spectra.py:351-364), which reads the lower triangle of the
// row-major slice (LAPACK ?heevd with UPLO='L').  Here one rocSOLVER divide-and-conquer call
// (?heevd, strided batched) decomposes all slices on the device:
//  * the slice is copied into the eigenvector output and decomposed there in place;
//  * rocSOLVER is column-major, so it sees the row-major slice C as C^T = conj(C); its upper
//    triangle is the row-major lower triangle numpy reads, its eigenvalues are those of C and
//    its eigenvectors are conj(v), stored as rows;
//  * a conjugate-transpose kernel turns that into numpy's layout (eigenvector e in column e).
// Eigenvalues come out in ascending order, as numpy's.  rocSOLVER's scratch memory is the
// caller's workspace (rocblas_set_workspace), so no device memory is allocated here.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>

#include <string>

#include "../../include/abf.h"
#include "abf_status.h"

namespace {

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// in-place conjugate transpose of each [M][M] slice; the thread of (i, j), i < j, swaps the
// pair, the diagonal thread conjugates
template <typename C>
__global__ __launch_bounds__(256) void conj_transpose_kernel(C *__restrict__ a, int M)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * M) return;
    const int i = (int)(idx / M), j = (int)(idx % M);
    if (i > j) return;
    C *m = a + (int64_t)blockIdx.y * M * M;
    const C x = m[(int64_t)i * M + j];
    if (i == j) {
        m[(int64_t)i * M + j] = C{x.x, -x.y};
        return;
    }
    const C y = m[(int64_t)j * M + i];
    m[(int64_t)i * M + j] = C{y.x, -y.y};
    m[(int64_t)j * M + i] = C{x.x, -x.y};
}

__global__ __launch_bounds__(256) void widen_kernel(const float *__restrict__ src, double *__restrict__ dst, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

// one rocBLAS handle per host thread (kept for the life of the process)
rocblas_handle handle()
{
    thread_local rocblas_handle h = nullptr;
    if (!h && rocblas_create_handle(&h) != rocblas_status_success) h = nullptr;
    return h;
}

struct EigLayout {
    size_t info, d, e, solver, total;
};

// [info int32 F][D real F*M][E real F*M][rocSOLVER scratch]
EigLayout layout(int F, int M, bool c128, size_t solver_bytes)
{
    const size_t real = c128 ? sizeof(double) : sizeof(float);
    EigLayout l;
    l.info = 0;
    l.d = align256((size_t)F * sizeof(int32_t));
    l.e = l.d + align256((size_t)F * M * real);
    l.solver = l.e + align256((size_t)F * M * real);
    l.total = l.solver + align256(solver_bytes);
    return l;
}

rocblas_status heevd(rocblas_handle h, bool c128, int M, void *A, void *D, void *E, int32_t *info, int F)
{
    const rocblas_stride sa = (rocblas_stride)M * M, sd = M;
    if (c128)
        return rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, M,
                                                static_cast<rocblas_double_complex *>(A), M, sa,
                                                static_cast<double *>(D), sd, static_cast<double *>(E), sd, info, F);
    return rocsolver_cheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, M,
                                            static_cast<rocblas_float_complex *>(A), M, sa, static_cast<float *>(D),
                                            sd, static_cast<float *>(E), sd, info, F);
}

// rocSOLVER scratch bytes of one call shape (a size query: nothing is launched or read)
bool solver_bytes(int F, int M, bool c128, size_t &bytes)
{
    rocblas_handle h = handle();
    if (!h) return false;
    if (rocblas_start_device_memory_size_query(h) != rocblas_status_success) return false;
    // non-null placeholders: arguments are checked for null but not accessed during a query
    alignas(16) static char dummy[64];
    (void)heevd(h, c128, M, dummy, dummy, dummy, reinterpret_cast<int32_t *>(dummy), F);
    return rocblas_stop_device_memory_size_query(h, &bytes) == rocblas_status_success;
}

} // namespace

extern "C" {

size_t abf_csm_eig_workspace_size(int32_t num_freqs, int32_t num_mics, int32_t is_c128)
{
    if (num_freqs <= 0 || num_mics <= 0) return 0;
    size_t sb = 0;
    if (!solver_bytes(num_freqs, num_mics, is_c128 != 0, sb)) return 0;
    return layout(num_freqs, num_mics, is_c128 != 0, sb).total;
}

int abf_csm_eig(const void *csm, int32_t num_freqs, int32_t num_mics, int32_t is_c128, double *eigval, void *eigvec,
                int32_t *info, void *workspace, size_t workspace_bytes, void *stream)
{
    if (num_freqs < 0 || num_mics < 0) return abf::set_error(ABF_ERR_INVALID, "bad sizes");
    if (num_freqs == 0 || num_mics == 0) return ABF_OK;
    if (!csm || !eigval || !eigvec) return abf::set_error(ABF_ERR_INVALID, "null pointer");
    const bool c128 = is_c128 != 0;
    const int F = num_freqs, M = num_mics;
    size_t sb = 0;
    if (!solver_bytes(F, M, c128, sb)) return abf::set_error(ABF_ERR_HIP, "rocSOLVER workspace query failed");
    const EigLayout l = layout(F, M, c128, sb);
    if (!workspace || workspace_bytes < l.total) return abf::set_error(ABF_ERR_WORKSPACE, "workspace too small");
    char *ws = static_cast<char *>(workspace);
    rocblas_handle h = handle();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!h || rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_workspace(h, sb ? ws + l.solver : nullptr, sb) != rocblas_status_success)
        return abf::set_error(ABF_ERR_HIP, "rocBLAS handle set-up failed");
    const size_t elem = c128 ? 16 : 8;
    if (hipMemcpyAsync(eigvec, csm, (size_t)F * M * M * elem, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return abf::set_error(ABF_ERR_HIP, "CSM copy failed");
    int32_t *inf = info ? info : reinterpret_cast<int32_t *>(ws + l.info);
    void *D = c128 ? static_cast<void *>(eigval) : static_cast<void *>(ws + l.d);
    const rocblas_status st = heevd(h, c128, M, eigvec, D, ws + l.e, inf, F);
    if (st != rocblas_status_success)
        return abf::set_error(ABF_ERR_HIP, ("rocSOLVER heevd failed with status " + std::to_string((int)st)).c_str());
    const dim3 grid((unsigned)(((int64_t)M * M + 255) / 256), (unsigned)F);
    if (c128)
        hipLaunchKernelGGL(conj_transpose_kernel<double2>, grid, dim3(256), 0, s, static_cast<double2 *>(eigvec), M);
    else
        hipLaunchKernelGGL(conj_transpose_kernel<float2>, grid, dim3(256), 0, s, static_cast<float2 *>(eigvec), M);
    if (!c128) {
        const int64_t n = (int64_t)F * M;
        hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<const float *>(ws + l.d), eigval, n);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return abf::set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
