// FP64 VALU issue-rate probe (no counterpart in the reference): the rate at
// which this device sustains independent FP64 FMA and multiply streams at
// the clock it runs under such a load.  The pair kernel's roofline reports
// its own VALU instruction rate against this measured ceiling beside the
// published FP64 peak, which assumes the peak engine clock.
#include <hip/hip_runtime.h>

#include "context.h"

namespace sdamd {
namespace {

constexpr int kChains = 8;

// each lane runs kChains independent chains of `iters` operations; the
// result is stored so that nothing is optimised away
template <bool Fma>
__global__ __launch_bounds__(256) void k_fp64_stream(int iters, double a, double b, double *__restrict__ out)
{
    double c[kChains];
    for (int k = 0; k < kChains; ++k) c[k] = (double)threadIdx.x * 1e-3 + (double)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) c[k] = Fma ? __builtin_fma(c[k], a, b) : c[k] * a;
    }
    double s = 0.0;
    for (int k = 0; k < kChains; ++k) s += c[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace
}  // namespace sdamd

extern "C" int sdddst_fp64_probe(int iters, double *fma_tflops, double *fma_valu_per_s, double *mul_valu_per_s)
{
    using namespace sdamd;
    SD_CHECK(iters > 0 && iters <= (1 << 20), SDDDST_ERR_ARG, "iters %d outside [1, 2^20]", iters);
    int dev = 0, cus = 0;
    SD_HIP(hipGetDevice(&dev));
    SD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 16 waves of 64 lanes per CU: enough to fill every SIMD
    const int threads = 256, blocks = cus * 4;
    double *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SDDDST_OK;
    double best[2] = {0.0, 0.0};
    if (hipMalloc(&out, sizeof(double) * (size_t)blocks * threads) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("fp64 probe: device allocation failed");
        rc = SDDDST_ERR_HIP;
    }
    // per kind: three timed launches after one to raise the clock; the best
    for (int kind = 0; kind < 2 && rc == SDDDST_OK; ++kind) {
        for (int rep = 0; rep < 4 && rc == SDDDST_OK; ++rep) {
            if (hipEventRecord(e0, st) != hipSuccess) {
                set_error("fp64 probe: event record failed");
                rc = SDDDST_ERR_HIP;
                break;
            }
            if (kind == 0)
                hipLaunchKernelGGL(k_fp64_stream<true>, dim3(blocks), dim3(threads), 0, st, iters, 0.999999, 1e-9, out);
            else
                hipLaunchKernelGGL(k_fp64_stream<false>, dim3(blocks), dim3(threads), 0, st, iters, 0.9999999, 0.0,
                                   out);
            float ms = 0.0f;
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
                set_error("fp64 probe: launch or timing failed");
                rc = SDDDST_ERR_HIP;
                break;
            }
            const double insts = (double)blocks * threads / 64.0 * iters * kChains;
            if (rep > 0 && insts / (ms * 1e-3) > best[kind]) best[kind] = insts / (ms * 1e-3);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    if (rc) return rc;
    if (fma_valu_per_s) *fma_valu_per_s = best[0];
    if (mul_valu_per_s) *mul_valu_per_s = best[1];
    if (fma_tflops) *fma_tflops = best[0] * 64.0 * 2.0 / 1e12;
    return SDDDST_OK;
}
