// Strong-row block preconditioner for FGMRES (no counterpart in the
// reference, which factors A directly with UMFPACK).
//
// At a step size far above what the configuration allows (the first attempt
// of a large run with point defects, h = 1e-6 at N = 65536) A = I + (1+d)/2 M
// is indefinite and ill-conditioned (condition ~1e6): the pair blocks do not
// precondition it and FGMRES stagnates.  The trouble sits in the rows whose
// diagonal is far from 1 or that hold a large off-diagonal entry (close
// pairs, dislocations pinned by point defects); the rest is a small
// perturbation of the identity.  So M^-1 solves the block A_HH of those rows
// H exactly -- a dense LU of |H| x |H| (rocSOLVER), a fraction of the N x N
// factorisation of the direct solver -- and applies the pair-block inverse on
// the other rows.  FGMRES is flexible, so this mixed inverse is a valid right
// preconditioner.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "context.h"
#include "solvers.h"

namespace sdamd {

namespace {

// flag[i] = |A_ii - 1| > thr or max_{j != i} |A_ij| > thr, from the CSR view
// (the pattern is symmetric: row i's column indices are column i's rows).
// One wavefront per row.
__global__ __launch_bounds__(256) void k_strong_mark(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                     const int *__restrict__ diag, const double *__restrict__ AxRow,
                                                     double thr, int *__restrict__ flag)
{
    const int lane = (int)(threadIdx.x & 63);
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const int dpos = diag[i];
    double mx = 0.0;
    for (int k = Ap[i] + lane; k < Ap[i + 1]; k += 64)
        if (k != dpos) mx = fmax(mx, fabs(AxRow[k]));
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if (lane == 0) flag[i] = (fabs(AxRow[dpos] - 1.0) > thr || mx > thr) ? 1 : 0;
}

__global__ void k_strong_zero(double *__restrict__ p, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) p[k] = 0.0;
}

// A_HH into the column-major nh x nh matrix D: element (pos[i], pos[j]) of
// each flagged row i and flagged column j.  One wavefront per row.
__global__ __launch_bounds__(256) void k_strong_gather(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                       const double *__restrict__ AxRow, const int *__restrict__ flag,
                                                       const int *__restrict__ pos, int nh, double *__restrict__ D)
{
    const int lane = (int)(threadIdx.x & 63);
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n || !flag[i]) return;
    const size_t r = (size_t)pos[i];
    for (int k = Ap[i] + lane; k < Ap[i + 1]; k += 64) {
        const int j = Ai[k];
        if (flag[j]) D[(size_t)pos[j] * (size_t)nh + r] = AxRow[k];
    }
}

__global__ void k_strong_get(int n, const int *__restrict__ flag, const int *__restrict__ pos,
                             const double *__restrict__ v, double *__restrict__ hv)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n && flag[i]) hv[pos[i]] = v[i];
}

__global__ void k_strong_put(int n, const int *__restrict__ flag, const int *__restrict__ pos,
                             const double *__restrict__ hv, double *__restrict__ z)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n && flag[i]) z[i] = hv[pos[i]];
}

inline unsigned blocks256(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int StrongBlock::factor(Context &c, double thr)
{
    const int n = c.n;
    if (!flag || n_alloc != n) {
        release_rows();
        SD_HIP(hipMalloc(&flag, sizeof(int) * n));
        SD_HIP(hipMalloc(&pos, sizeof(int) * (n + 1)));
        SD_HIP(hipMalloc(&total, sizeof(long long)));
        SD_HIP(hipMalloc(&hv, sizeof(double) * n));
        n_alloc = n;
    }
    if (!handle) {
        if (rocblas_create_handle(&handle) != rocblas_status_success) {
            handle = nullptr;
            set_error("rocblas_create_handle failed");
            return SDDDST_ERR_SOLVER;
        }
        if (!info) SD_HIP(hipMalloc(&info, sizeof(int64_t)));
    }
    rocblas_set_stream(handle, c.stream);
    hipLaunchKernelGGL(k_strong_mark, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.AxRow, thr,
                       flag);
    SD_HIP(hipGetLastError());
    int rc = launch_scan(c, n, flag, pos, total);
    if (rc) return rc;
    long long count = 0;
    if ((rc = host_wait(c, total, 1, &count))) return rc;
    nh = (int)count;
    if (nh == 0) return SDDDST_OK;
    SD_CHECK(nh <= kMaxRows, SDDDST_ERR_SOLVER,
             "strong-row block: %d rows exceed the limit %d of its dense factorisation", nh, kMaxRows);
    if (nh > nh_alloc) {
        if (D) (void)hipFree(D);
        if (piv) (void)hipFree(piv);
        D = nullptr;
        piv = nullptr;
        const size_t bytes = sizeof(double) * (size_t)nh * (size_t)nh;
        if (hipMalloc(&D, bytes) != hipSuccess) {
            D = nullptr;
            nh_alloc = 0;
            set_error("strong-row block: cannot allocate %zu bytes", bytes);
            return SDDDST_ERR_NOMEM;
        }
        SD_HIP(hipMalloc(&piv, sizeof(int64_t) * nh));
        nh_alloc = nh;
    }
    hipLaunchKernelGGL(k_strong_zero, dim3(2048), dim3(256), 0, c.stream, D, (size_t)nh * (size_t)nh);
    hipLaunchKernelGGL(k_strong_gather, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.AxRow, flag, pos,
                       nh, D);
    SD_HIP(hipGetLastError());
    if (rocsolver_dgetrf_64(handle, nh, nh, D, nh, piv, info) != rocblas_status_success) {
        set_error("strong-row block: rocsolver_dgetrf_64 failed");
        return SDDDST_ERR_SOLVER;
    }
    int64_t hinfo = 0;
    if ((rc = host_wait(c, info, 1, &hinfo))) return rc;
    SD_CHECK(hinfo == 0, SDDDST_ERR_SOLVER, "strong-row block: A_HH is singular (U(%lld,%lld) = 0)",
             (long long)hinfo, (long long)hinfo);
    return SDDDST_OK;
}

int StrongBlock::apply(Context &c, const double *v, double *z)
{
    // z already holds the pair-block inverse applied to v; the rows of H are
    // replaced by A_HH^-1 v_H
    if (nh == 0) return SDDDST_OK;
    const int n = c.n;
    hipLaunchKernelGGL(k_strong_get, dim3(blocks256(n)), dim3(256), 0, c.stream, n, flag, pos, v, hv);
    SD_HIP(hipGetLastError());
    if (rocsolver_dgetrs_64(handle, rocblas_operation_none, nh, 1, D, nh, piv, hv, nh) != rocblas_status_success) {
        set_error("strong-row block: rocsolver_dgetrs_64 failed");
        return SDDDST_ERR_SOLVER;
    }
    hipLaunchKernelGGL(k_strong_put, dim3(blocks256(n)), dim3(256), 0, c.stream, n, flag, pos, hv, z);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

void StrongBlock::release_rows()
{
    for (void *p : {(void *)flag, (void *)pos, (void *)total, (void *)hv})
        if (p) (void)hipFree(p);
    flag = pos = nullptr;
    total = nullptr;
    hv = nullptr;
    n_alloc = 0;
}

void StrongBlock::release()
{
    release_rows();
    for (void *p : {(void *)D, (void *)piv, (void *)info})
        if (p) (void)hipFree(p);
    D = nullptr;
    piv = info = nullptr;
    nh = nh_alloc = 0;
    if (handle) rocblas_destroy_handle(handle);
    handle = nullptr;
}

}  // namespace sdamd
