// Frequency transforms of libwlsq_hip: the reference's MKL FFT wrappers
// (cpp_src/mkl_fft.cpp:9-39) and the shot FFT of the driver
// (main.py:119-123, scipy.fftpack.fft along time), on the device with
// hipFFT (rocFFT). Plans are cached per (device, shape, strides, stream).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "wlsq_internal.h"

using namespace wlsq;

namespace {

// rows j in (nt/2, nt) of a real signal's spectrum: X[j] = conj(X[nt - j])
__global__ void mirror_rows_kernel(float2 *out, int ld, int nt, int nx)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx)
        return;
    for (int j = nt / 2 + 1 + blockIdx.y; j < nt; j += gridDim.y) {
        const float2 v = out[(long long)(nt - j) * ld + x];
        out[(long long)j * ld + x] = make_float2(v.x, -v.y);
    }
}

// spectrum of a length-1 signal: the sample itself
__global__ void real_to_complex_row_kernel(const float *in, float2 *out, int nx)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < nx)
        out[x] = make_float2(in[x], 0.f);
}

__global__ void scale_kernel(float2 *p, int n, float s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = make_float2(p[i].x * s, p[i].y * s);
}

const char *fft_error_name(hipfftResult r)
{
    switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_INVALID_PLAN: return "invalid plan";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INTERNAL_ERROR: return "internal error";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    case HIPFFT_NOT_SUPPORTED: return "not supported";
    default: return "hipFFT error";
    }
}

#define FFT_TRY(expr)                                                                          \
    do {                                                                                       \
        hipfftResult r_ = (expr);                                                              \
        if (r_ != HIPFFT_SUCCESS)                                                              \
            return set_error(r_ == HIPFFT_ALLOC_FAILED ? WLSQ_ENOMEM : WLSQ_EHIP, "%s: %s",    \
                             #expr, fft_error_name(r_));                                       \
    } while (0)

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(e_ == hipErrorOutOfMemory ? WLSQ_ENOMEM : WLSQ_EHIP, "%s: %s",    \
                             #expr, hipGetErrorString(e_));                                    \
    } while (0)

// (device, kind, n, batch, in stride, out stride, stream) -> plan
using PlanKey = std::tuple<int, int, int, int, int, int, hipStream_t>;
std::mutex g_fft_mu;  // plan cache, host-entry stream and staging buffer
std::map<PlanKey, hipfftHandle> g_plans;
std::map<int, hipStream_t> g_host_streams;
std::map<int, std::pair<void *, size_t>> g_stage;

int plan_for(const PlanKey &key, hipfftHandle *out)
{
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        *out = it->second;
        return WLSQ_OK;
    }
    const int kind = std::get<1>(key), n = std::get<2>(key), batch = std::get<3>(key);
    const int ist = std::get<4>(key), ost = std::get<5>(key);
    hipfftHandle p;
    if (kind == 0) {  // complex, contiguous
        FFT_TRY(hipfftPlan1d(&p, n, HIPFFT_C2C, 1));
    } else {  // real columns of stride ist -> complex columns of stride ost
        int nn[1] = {n};
        int inembed[1] = {n};
        int onembed[1] = {n / 2 + 1};
        FFT_TRY(hipfftPlanMany(&p, 1, nn, inembed, ist, 1, onembed, ost, 1, HIPFFT_R2C, batch));
    }
    hipfftResult r = hipfftSetStream(p, std::get<6>(key));
    if (r != HIPFFT_SUCCESS) {
        hipfftDestroy(p);
        return set_error(WLSQ_EHIP, "hipfftSetStream: %s", fft_error_name(r));
    }
    g_plans[key] = p;
    *out = p;
    return WLSQ_OK;
}

// In-place transform of a host array on the device's own stream.
int c64fft1d(wlsq_c32 *data, int N, bool forward)
{
    if (N < 0)
        return set_error(WLSQ_EINVAL, "negative length %d", N);
    if (N == 0)
        return WLSQ_OK;
    if (!data)
        return set_error(WLSQ_EINVAL, "NULL array");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_fft_mu);
    hipStream_t &st = g_host_streams[dev];
    if (!st)
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    auto &stage = g_stage[dev];
    const size_t bytes = sizeof(float2) * (size_t)N;
    if (stage.second < bytes) {
        if (stage.first)
            HIP_TRY(hipFree(stage.first));
        stage.first = nullptr;
        stage.second = 0;
        HIP_TRY(hipMalloc(&stage.first, bytes));
        stage.second = bytes;
    }
    hipfftHandle p;
    int rc = plan_for(PlanKey{dev, 0, N, 1, 1, 1, st}, &p);
    if (rc != WLSQ_OK)
        return rc;
    float2 *buf = static_cast<float2 *>(stage.first);
    HIP_TRY(hipMemcpyAsync(buf, data, bytes, hipMemcpyHostToDevice, st));
    FFT_TRY(hipfftExecC2C(p, reinterpret_cast<hipfftComplex *>(buf),
                          reinterpret_cast<hipfftComplex *>(buf),
                          forward ? HIPFFT_FORWARD : HIPFFT_BACKWARD));
    if (!forward) {
        hipLaunchKernelGGL(scale_kernel, dim3((N + 255) / 256), dim3(256), 0, st, buf, N,
                           1.0f / (float)N);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(data, buf, bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WLSQ_OK;
}

}  // namespace

extern "C" {

int c64fft1dforw(wlsq_c32 *data, int N) { return c64fft1d(data, N, true); }

int c64fft1dback(wlsq_c32 *data, int N) { return c64fft1d(data, N, false); }

int wlsq_trace_spectra_device(int nt, int nx, const float *traces, int ld_in, wlsq_c32 *out,
                              int ld_out, void *stream)
{
    if (nt < 0 || nx < 0)
        return set_error(WLSQ_EINVAL, "negative size (nt=%d nx=%d)", nt, nx);
    if (nt == 0 || nx == 0)
        return WLSQ_OK;
    if (!traces || !out)
        return set_error(WLSQ_EINVAL, "NULL array");
    if (ld_in < nx || ld_out < nx)
        return set_error(WLSQ_EINVAL, "row stride below nx (ld_in=%d ld_out=%d nx=%d)", ld_in,
                         ld_out, nx);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nt == 1) {
        hipLaunchKernelGGL(real_to_complex_row_kernel, dim3((nx + 255) / 256), dim3(256), 0, st,
                           traces, reinterpret_cast<float2 *>(out), nx);
        HIP_TRY(hipGetLastError());
        return WLSQ_OK;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_fft_mu);
    hipfftHandle p;
    int rc = plan_for(PlanKey{dev, 1, nt, nx, ld_in, ld_out, st}, &p);
    if (rc != WLSQ_OK)
        return rc;
    FFT_TRY(hipfftExecR2C(p, const_cast<hipfftReal *>(traces), reinterpret_cast<hipfftComplex *>(out)));
    const int upper = nt - (nt / 2 + 1);  // rows filled by symmetry
    if (upper > 0) {
        const int gy = upper < 1024 ? upper : 1024;
        hipLaunchKernelGGL(mirror_rows_kernel, dim3((nx + 255) / 256, gy), dim3(256), 0, st,
                           reinterpret_cast<float2 *>(out), ld_out, nt, nx);
        HIP_TRY(hipGetLastError());
    }
    return WLSQ_OK;
}

}  // extern "C"
