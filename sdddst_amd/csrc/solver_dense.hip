// Device direct solver for the Newton correction A delta = g: LU with partial
// pivoting of the assembled Jacobian (rocSOLVER getrf / getrs).  It stands in
// for the host UMFPACK factorisation (simulation.cpp:386-396) when no
// libumfpack is installed, with the same call pattern: one factorisation per
// integrate, ic solves against it.  Dense storage is N^2 doubles
// (2 GiB at N = 16384, 32 GiB at N = 65536: within one MI355X's HBM).  The
// 64-bit (ILP64) rocSOLVER interfaces are used: above N = 46340 the matrix
// has more than 2^31 elements, beyond the 32-bit interface's index range.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "context.h"
#include "solvers.h"

namespace sdamd {

namespace {

// Column j of the CSC into column j of the column-major dense matrix.
__global__ __launch_bounds__(256) void k_scatter_dense(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                       const double *__restrict__ Ax, double *__restrict__ A)
{
    const int lane = (int)(threadIdx.x & 63);
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    double *col = A + (size_t)j * n;
    for (int k = Ap[j] + lane; k < Ap[j + 1]; k += 64) col[Ai[k]] = Ax[k];
}

__global__ void k_zero(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = 0.0;
}

}  // namespace

int DenseLu::init(Context &c)
{
    if (A) return SDDDST_OK;
    const size_t bytes = sizeof(double) * (size_t)c.n * c.n;
    if (hipMalloc(&A, bytes) != hipSuccess) {
        A = nullptr;
        set_error("dense LU: cannot allocate %zu bytes", bytes);
        return SDDDST_ERR_NOMEM;
    }
    SD_HIP(hipMalloc(&ipiv, sizeof(int64_t) * c.n));
    SD_HIP(hipMalloc(&info, sizeof(int64_t)));
    if (rocblas_create_handle(&handle) != rocblas_status_success) {
        set_error("rocblas_create_handle failed");
        return SDDDST_ERR_SOLVER;
    }
    rocblas_set_stream(handle, c.stream);
    return SDDDST_OK;
}

void DenseLu::release()
{
    if (A) (void)hipFree(A);
    if (ipiv) (void)hipFree(ipiv);
    if (info) (void)hipFree(info);
    if (handle) rocblas_destroy_handle(handle);
    A = nullptr;
    ipiv = nullptr;
    info = nullptr;
    handle = nullptr;
}

int DenseLu::factor(Context &c)
{
    int rc = init(c);
    if (rc) return rc;
    const int n = c.n;
    const size_t total = (size_t)n * n;
    if ((rc = ensure_ax(c))) return rc;
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, c.stream, A, total);
    SD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scatter_dense, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.Ax, A);
    SD_HIP(hipGetLastError());
    if (rocsolver_dgetrf_64(handle, n, n, A, n, ipiv, info) != rocblas_status_success) {
        set_error("rocsolver_dgetrf_64 failed");
        return SDDDST_ERR_SOLVER;
    }
    int64_t hinfo = 0;
    SD_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    SD_CHECK(hinfo == 0, SDDDST_ERR_SOLVER, "dense LU: matrix is singular (U(%lld,%lld) = 0)", (long long)hinfo,
             (long long)hinfo);
    return SDDDST_OK;
}

int DenseLu::solve(Context &c)
{
    SD_HIP(hipMemcpyAsync(c.delta, c.g, sizeof(double) * c.n, hipMemcpyDeviceToDevice, c.stream));
    if (rocsolver_dgetrs_64(handle, rocblas_operation_none, c.n, 1, A, c.n, ipiv, c.delta, c.n) !=
        rocblas_status_success) {
        set_error("rocsolver_dgetrs_64 failed");
        return SDDDST_ERR_SOLVER;
    }
    return SDDDST_OK;
}

}  // namespace sdamd
