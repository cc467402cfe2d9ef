// Fourier resampling of one float64 signal (SignalGenerator.usignal, acoular/signals.py:93-126,
// which calls scipy.signal.resample) for gfx950: the one-sided spectrum of the n input samples,
// truncated or zero-padded to num / 2 + 1 bins, the Nyquist bin of an even length halved when
// upsampling (scipy's split of that component between +N/2 and -N/2), the inverse transform of
// length num and the factor num / n of scipy's normalisation.  The transforms are double-precision
// rocFFT plans (lengths like 51,200 = 2^11 5^2 are not powers of two), kept per length for the
// life of the process; the spectrum step and the scaling are kernels of this file.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <utility>

#include "abf_status.h"
#include "../../include/abf.h"

namespace abf {

// Y[k] = X[k] for k < nyq, 0 above; Y[half] halved (half < 0: none)
__global__ __launch_bounds__(256) void resample_spectrum_kernel(const double2 *__restrict__ X, double2 *__restrict__ Y,
                                                                int64_t ny, int64_t nyq, int64_t half)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ny) return;
    double2 v = k < nyq ? X[k] : make_double2(0.0, 0.0);
    if (k == half) v = make_double2(v.x * 0.5, v.y * 0.5);
    Y[k] = v;
}

__global__ __launch_bounds__(256) void scale_f64_kernel(double *__restrict__ y, int64_t n, double s)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] *= s;
}

namespace {

std::mutex g_fft_mu;
std::map<std::pair<int64_t, int>, std::pair<rocfft_plan, size_t>> g_fft_plans; // (length, forward) -> plan, work

int fft_plan(int64_t n, bool forward, rocfft_plan &plan, size_t &work)
{
    std::lock_guard<std::mutex> lk(g_fft_mu);
    static bool setup = false;
    if (!setup) {
        if (rocfft_setup() != rocfft_status_success) return set_error(ABF_ERR_HIP, "rocfft_setup failed");
        setup = true;
    }
    const auto key = std::make_pair(n, forward ? 1 : 0);
    auto it = g_fft_plans.find(key);
    if (it == g_fft_plans.end()) {
        rocfft_plan p = nullptr;
        const size_t len = (size_t)n;
        if (rocfft_plan_create(&p, rocfft_placement_notinplace,
                               forward ? rocfft_transform_type_real_forward : rocfft_transform_type_real_inverse,
                               rocfft_precision_double, 1, &len, 1, nullptr) != rocfft_status_success)
            return set_error(ABF_ERR_HIP, "rocfft_plan_create failed");
        size_t w = 0;
        if (rocfft_plan_get_work_buffer_size(p, &w) != rocfft_status_success)
            return set_error(ABF_ERR_HIP, "rocfft_plan_get_work_buffer_size failed");
        it = g_fft_plans.emplace(key, std::make_pair(p, w)).first;
    }
    plan = it->second.first;
    work = it->second.second;
    return ABF_OK;
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace
} // namespace abf

using namespace abf;

extern "C" {

size_t abf_resample_workspace_size(int64_t num_samples, int64_t num)
{
    if (num_samples < 1 || num < num_samples) return 0;
    rocfft_plan pf, pi;
    size_t wf = 0, wi = 0;
    if (fft_plan(num_samples, true, pf, wf) || fft_plan(num, false, pi, wi)) return 0;
    return up256((size_t)num_samples * sizeof(double)) + up256((size_t)(num_samples / 2 + 1) * sizeof(double2)) +
           up256((size_t)(num / 2 + 1) * sizeof(double2)) + up256(wf > wi ? wf : wi);
}

int abf_resample(const double *x, int64_t num_samples, int64_t num, double *y, void *workspace,
                 size_t workspace_bytes, void *stream)
{
    const int64_t n = num_samples;
    if (n < 1 || num < n) return set_error(ABF_ERR_UNSUPPORTED, "resampling supports num >= num_samples >= 1");
    if (!x || !y) return set_error(ABF_ERR_INVALID, "null pointer");
    rocfft_plan pf, pi;
    size_t wf = 0, wi = 0;
    if (int r = fft_plan(n, true, pf, wf)) return r;
    if (int r = fft_plan(num, false, pi, wi)) return r;
    const size_t need = abf_resample_workspace_size(n, num);
    if (!workspace || workspace_bytes < need) return set_error(ABF_ERR_WORKSPACE, "workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    double *xin = reinterpret_cast<double *>(ws);
    double2 *X = reinterpret_cast<double2 *>(ws + up256((size_t)n * sizeof(double)));
    double2 *Y = reinterpret_cast<double2 *>(reinterpret_cast<char *>(X) + up256((size_t)(n / 2 + 1) * sizeof(double2)));
    void *work = reinterpret_cast<char *>(Y) + up256((size_t)(num / 2 + 1) * sizeof(double2));
    // a not-in-place real transform may overwrite its input: transform a copy
    if (hipMemcpyAsync(xin, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return set_error(ABF_ERR_HIP, "copy before the forward transform failed");
    rocfft_execution_info info = nullptr;
    if (rocfft_execution_info_create(&info) != rocfft_status_success)
        return set_error(ABF_ERR_HIP, "rocfft_execution_info_create failed");
    int rc = ABF_OK;
    if (rocfft_execution_info_set_stream(info, s) != rocfft_status_success) rc = ABF_ERR_HIP;
    const size_t wmax = wf > wi ? wf : wi;
    if (!rc && wmax && rocfft_execution_info_set_work_buffer(info, work, wmax) != rocfft_status_success)
        rc = ABF_ERR_HIP;
    void *in_f[1] = {xin}, *out_f[1] = {X};
    if (!rc && rocfft_execute(pf, in_f, out_f, info) != rocfft_status_success) rc = ABF_ERR_HIP;
    if (!rc) {
        const int64_t N = n; // min(num, n) for num >= n
        const int64_t nyq = N / 2 + 1, ny = num / 2 + 1;
        const int64_t half = (N % 2 == 0 && n < num) ? N / 2 : -1;
        hipLaunchKernelGGL(resample_spectrum_kernel, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, s, X, Y, ny,
                           nyq, half);
        void *in_i[1] = {Y}, *out_i[1] = {y};
        if (rocfft_execute(pi, in_i, out_i, info) != rocfft_status_success) rc = ABF_ERR_HIP;
    }
    if (!rc)
        hipLaunchKernelGGL(scale_f64_kernel, dim3((unsigned)((num + 255) / 256)), dim3(256), 0, s, y, num,
                           1.0 / (double)n);
    rocfft_execution_info_destroy(info);
    if (rc) return set_error(rc, "rocFFT execution failed");
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ABF_ERR_HIP, hipGetErrorString(e));
    return ABF_OK;
}

} // extern "C"
