// Array evaluation of the device field and elementary functions
// (sdddst_field_eval / sdddst_libm_eval).
#include <hip/hip_runtime.h>

#include "context.h"
#include "field.h"

namespace sdamd {
namespace {

__global__ void k_field_eval(int which, int64_t n, const double *__restrict__ dx, const double *__restrict__ dy,
                             double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r;
    if (which == 0) r = sdfield::xy(dx[i], dy[i]);
    else if (which == 1) r = sdfield::xy_diff_x(dx[i], dy[i]);
    else if (which == 2) r = sdfield::xy_fast(dx[i], dy[i]);
    else r = sdfield::xy_diff_x2(dx[i], dy[i]);
    out[i] = r;
}

__global__ void k_libm_eval(int fn, int64_t n, const double *__restrict__ x, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r = 0;
    switch (fn) {
    case 0: r = sdlibm::cos(v); break;
    case 1: r = sdlibm::sin(v); break;
    case 2: r = sdlibm::exp(v); break;
    case 3: r = sdlibm::expm1(v); break;
    case 4: r = sdlibm::cosh(v); break;
    case 5: r = sdlibm::sinh(v); break;
    case 6: r = sdlibm::sqrt(v); break;
    case 7: r = sdlibm::pow_one_third(v); break;
    case 8: r = sdfield::normalize(v); break;
    default: r = 0; break;
    }
    out[i] = r;
}

}  // namespace

int field_eval(Context &c, int which, int64_t n, const double *dx, const double *dy, double *out)
{
    SD_CHECK(which >= 0 && which <= 3 && n >= 0 && dx && dy && out, SDDDST_ERR_ARG, "bad arguments");
    if (n == 0) return SDDDST_OK;
    double *d = nullptr;
    SD_HIP(hipMalloc(&d, sizeof(double) * 3 * n));
    SD_HIP(hipMemcpyAsync(d, dx, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipMemcpyAsync(d + n, dy, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_field_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, which, n, d, d + n,
                       d + 2 * n);
    SD_HIP(hipGetLastError());
    SD_HIP(hipMemcpyAsync(out, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    SD_HIP(hipFree(d));
    return SDDDST_OK;
}

int libm_eval(Context &c, int fn, int64_t n, const double *x, double *out)
{
    SD_CHECK(fn >= 0 && fn <= 8 && n >= 0 && x && out, SDDDST_ERR_ARG, "bad arguments");
    if (n == 0) return SDDDST_OK;
    double *d = nullptr;
    SD_HIP(hipMalloc(&d, sizeof(double) * 2 * n));
    SD_HIP(hipMemcpyAsync(d, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_libm_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, fn, n, d, d + n);
    SD_HIP(hipGetLastError());
    SD_HIP(hipMemcpyAsync(out, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    SD_HIP(hipFree(d));
    return SDDDST_OK;
}

}  // namespace sdamd
