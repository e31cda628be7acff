// EV analysis (next to the step path): the dense Hessian of the dislocation
// energy, its eigen-decomposition and the per-mode observables that the
// reference's time-series processor writes.
//   Hessian        SpectralDecompositor::decompose (spectral_decompositor.cpp:
//                  96-160): off the diagonal b_i b_j dtau/dx (x_i - x_j,
//                  y_i - y_j) for i > j, mirrored; on it minus the lower
//                  row's sum, minus the point-defect terms (A = 1.6e-3 /
//                  sqrt(defects), K = 1.65^2 1e6 / 256 N), minus the column
//                  below, in that order.
//   eigensolve     LAPACKE_dsyevd (spectral_decompositor.cpp:180-191):
//                  rocSOLVER dsyevd, ascending eigenvalues, eigenvector i in
//                  row i of the row-major result (column i of the reference's
//                  column-major array).
//   observables    time_series_processor.cpp:40-58, 75-82: participation
//                  number 1 / sum_j v_ij^4, s_i = sum_j v_ij b_j and
//                  s_i^2 / lambda_i.
//   derivatives    time_series_processor.cpp:92-111, 205-246: with the
//                  third-derivative tensor T_ijk of calculateHessianDerivative
//                  and speeds s, the reference writes per mode
//                  sum_ij v_i v_j sum_k s_k T_ijk / sigma'.  The only nonzero
//                  T are T_iii = -b_i sum_l b_l X_il, T_iik = b_i b_k X_ik,
//                  T_iji = b_i b_j X_ij and T_ijj = -b_i b_j X_ij, with
//                  X_ij = xy_diff_x2(wrap(x_i - x_j), wrap(y_i - y_j)), so
//                  C_ij = sum_k s_k T_ijk = (s_i - s_j) b_i b_j X_ij off the
//                  diagonal and C_ii = -sum_{k != i} C_ik: the dense C, one
//                  dgemm W = C V and a dot per mode (O(n^3) where the
//                  reference's loop is O(n^4)).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "context.h"
#include "field.h"

namespace sdamd {

namespace {

// H[i][j] = H[j][i] = b_i b_j xy_diff_x(wrap(x_i - x_j), wrap(y_i - y_j)) for
// i > j; one thread per (i, j) of the strictly lower triangle.
__global__ void k_hessian_pairs(int n, const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ b, double *__restrict__ H)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (i >= n || j >= i) return;
    const double dx = sdfield::normalize(x[i] - x[j]);
    const double dy = sdfield::normalize(y[i] - y[j]);
    const double t = b[i] * b[j] * sdfield::xy_diff_x(dx, dy);
    H[(size_t)i * n + j] = t;
    H[(size_t)j * n + i] = t;
}

// The diagonal, one thread per row, in the reference's operation order.  The
// lower row's entries are read from column i (H[j][i] holds the same bits as
// H[i][j]), so neighbouring threads read neighbouring addresses.
__global__ void k_hessian_diag(int n, int pc, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ b, const double *__restrict__ px,
                               const double *__restrict__ py, double A, double K, double *__restrict__ H)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double subSum = 0;
    for (int j = 0; j < i; ++j) subSum += H[(size_t)j * n + i];
    double d = 0.0;
    d -= subSum;
    for (int l = 0; l < pc; ++l) {
        const double dx = sdfield::normalize(x[i] - px[l]);
        const double dy = sdfield::normalize(y[i] - py[l]);
        d -= b[i] * sdfield::pd_force_dx(dx, dy, A, K);
    }
    for (int j = i + 1; j < n; ++j) d -= H[(size_t)j * n + i];
    H[(size_t)i * n + i] = d;
}

// C_ij = (s_i - s_j) b_i b_j xy_diff_x2(wrap(x_i - x_j), wrap(y_i - y_j))
// for i != j; one thread per ordered pair, as the reference evaluates them.
__global__ void k_deriv_pairs(int n, const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ b, const double *__restrict__ s, double *__restrict__ C)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (i >= n || j >= n || i == j) return;
    const double dx = sdfield::normalize(x[i] - x[j]);
    const double dy = sdfield::normalize(y[i] - y[j]);
    C[(size_t)i * n + j] = (s[i] - s[j]) * (b[i] * b[j] * sdfield::xy_diff_x2(dx, dy));
}

// C_ii = -sum_{k != i} C_ik, one thread per row.
__global__ void k_deriv_diag(int n, double *__restrict__ C)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double *row = C + (size_t)i * n;
    double acc = 0.0;
    for (int k = 0; k < n; ++k)
        if (k != i) acc += row[k];
    C[(size_t)i * n + i] = -acc;
}

// deriv[m] = (sum_i W_mi V_mi) / rate with W = C V: row m of W (row-major)
// is C applied to eigenvector m.  One workgroup per mode.
__global__ __launch_bounds__(256) void k_deriv_modes(int n, const double *__restrict__ V,
                                                     const double *__restrict__ W, double rate,
                                                     double *__restrict__ deriv)
{
    __shared__ double r[256];
    const int mode = (int)blockIdx.x, t = (int)threadIdx.x;
    const double *v = V + (size_t)mode * n, *w = W + (size_t)mode * n;
    double a = 0.0;
    for (int j = t; j < n; j += 256) a += w[j] * v[j];
    r[t] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) r[t] += r[t + off];
        __syncthreads();
    }
    if (t == 0) deriv[mode] = r[0] / rate;
}

// Per mode (one workgroup each, fixed-order reductions): 1 / sum_j v^4,
// s = sum_j v_j b_j and s^2 / lambda.
__global__ __launch_bounds__(256) void k_mode_observables(int n, const double *__restrict__ V,
                                                          const double *__restrict__ lambda,
                                                          const double *__restrict__ b, double *__restrict__ pn,
                                                          double *__restrict__ sum_b, double *__restrict__ gan)
{
    __shared__ double r4[256], rb[256];
    const int mode = (int)blockIdx.x, t = (int)threadIdx.x;
    const double *v = V + (size_t)mode * n;
    double a4 = 0.0, ab = 0.0;
    for (int j = t; j < n; j += 256) {
        const double e = v[j];
        const double e2 = e * e;
        a4 += e2 * e2;
        ab += e * b[j];
    }
    r4[t] = a4;
    rb[t] = ab;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            r4[t] += r4[t + off];
            rb[t] += rb[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double s = rb[0];
        if (pn) pn[mode] = 1.0 / r4[0];
        if (sum_b) sum_b[mode] = s;
        if (gan) gan[mode] = s * s / lambda[mode];
    }
}

}  // namespace

int ev_decompose(Context &c, int slot, double *eigenvalues, double *eigenvectors, double *pn, double *gan,
                 double *sum_b, const double *speeds, double rate, double *deriv)
{
    const int n = c.n;
    const size_t nn = (size_t)n * n;
    const bool with_deriv = speeds != nullptr && deriv != nullptr;
    double *H = nullptr, *lam = nullptr, *E = nullptr, *obs = nullptr, *C = nullptr, *W = nullptr, *sp = nullptr;
    int *info = nullptr;
    rocblas_handle handle = nullptr;
    int rc = SDDDST_OK;
    // spectral_decompositor.cpp:81-90: the point-defect constants of the analysis
    const double A = c.pc > 0 ? 1e-4 * 16.0 * 1. / sqrt((double)c.pc) : 0.0;
    const double K = c.pc > 0 ? 1.65 * 1.65 * 1e6 / 256.0 * double(n) : 0.0;
    int hinfo = 0;
    if (hipMalloc(&H, sizeof(double) * nn) != hipSuccess) {
        H = nullptr;
        (void)hipGetLastError();
        set_error("EV analysis: cannot allocate the %d x %d Hessian", n, n);
        return SDDDST_ERR_NOMEM;
    }
    if (hipMalloc(&lam, sizeof(double) * n) != hipSuccess || hipMalloc(&E, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&obs, sizeof(double) * 4 * n) != hipSuccess || hipMalloc(&info, sizeof(int)) != hipSuccess ||
        (with_deriv && (hipMalloc(&C, sizeof(double) * nn) != hipSuccess ||
                        hipMalloc(&W, sizeof(double) * nn) != hipSuccess ||
                        hipMalloc(&sp, sizeof(double) * n) != hipSuccess))) {
        (void)hipGetLastError();
        set_error("EV analysis: device allocation failed");
        rc = SDDDST_ERR_NOMEM;
        goto done;
    }
    {
        const dim3 blk(32, 8), grd((n + 31) / 32, (n + 7) / 8);
        hipLaunchKernelGGL(k_hessian_pairs, grd, blk, 0, c.stream, n, c.x[slot], c.y, c.b, H);
        hipLaunchKernelGGL(k_hessian_diag, dim3((n + 127) / 128), dim3(128), 0, c.stream, n, c.pc, c.x[slot], c.y,
                           c.b, c.px, c.py, A, K, H);
        if (hipGetLastError() != hipSuccess) {
            set_error("EV analysis: Hessian launch failed");
            rc = SDDDST_ERR_HIP;
            goto done;
        }
    }
    if (rocblas_create_handle(&handle) != rocblas_status_success) {
        handle = nullptr;
        set_error("rocblas_create_handle failed");
        rc = SDDDST_ERR_SOLVER;
        goto done;
    }
    rocblas_set_stream(handle, c.stream);
    // the Hessian is symmetric, so its row-major array is its column-major one
    if (rocsolver_dsyevd(handle, rocblas_evect_original, rocblas_fill_upper, n, H, n, lam, E, info) !=
        rocblas_status_success) {
        set_error("rocsolver_dsyevd failed");
        rc = SDDDST_ERR_SOLVER;
        goto done;
    }
    hipLaunchKernelGGL(k_mode_observables, dim3(n), dim3(256), 0, c.stream, n, H, lam, c.b, obs, obs + n,
                       obs + 2 * n);
    if (hipGetLastError() != hipSuccess) {
        set_error("EV analysis: observables launch failed");
        rc = SDDDST_ERR_HIP;
        goto done;
    }
    if (with_deriv) {
        if (hipMemcpyAsync(sp, speeds, sizeof(double) * n, hipMemcpyHostToDevice, c.stream) != hipSuccess) {
            set_error("EV analysis: speeds upload failed");
            rc = SDDDST_ERR_HIP;
            goto done;
        }
        const dim3 blk(32, 8), grd((n + 31) / 32, (n + 7) / 8);
        hipLaunchKernelGGL(k_deriv_pairs, grd, blk, 0, c.stream, n, c.x[slot], c.y, c.b, sp, C);
        hipLaunchKernelGGL(k_deriv_diag, dim3((n + 127) / 128), dim3(128), 0, c.stream, n, C);
        if (hipGetLastError() != hipSuccess) {
            set_error("EV analysis: derivative launch failed");
            rc = SDDDST_ERR_HIP;
            goto done;
        }
        // column-major views: C is symmetric, H holds eigenvector m in column
        // m, so W = C H has C v_m in column m (row m of the row-major array)
        const double one = 1.0, zero = 0.0;
        if (rocblas_dgemm(handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, C, n, H, n, &zero,
                          W, n) != rocblas_status_success) {
            set_error("EV analysis: rocblas_dgemm failed");
            rc = SDDDST_ERR_SOLVER;
            goto done;
        }
        hipLaunchKernelGGL(k_deriv_modes, dim3(n), dim3(256), 0, c.stream, n, H, W, rate, obs + 3 * n);
        if (hipGetLastError() != hipSuccess) {
            set_error("EV analysis: derivative launch failed");
            rc = SDDDST_ERR_HIP;
            goto done;
        }
    }
    if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        (eigenvalues && hipMemcpyAsync(eigenvalues, lam, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream)) ||
        (eigenvectors && hipMemcpyAsync(eigenvectors, H, sizeof(double) * nn, hipMemcpyDeviceToHost, c.stream)) ||
        (pn && hipMemcpyAsync(pn, obs, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream)) ||
        (sum_b && hipMemcpyAsync(sum_b, obs + n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream)) ||
        (gan && hipMemcpyAsync(gan, obs + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream)) ||
        (with_deriv && hipMemcpyAsync(deriv, obs + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream)) ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        set_error("EV analysis: copy to the host failed");
        rc = SDDDST_ERR_HIP;
        goto done;
    }
    if (hinfo != 0) {
        set_error("EV analysis: dsyevd did not converge (info %d)", hinfo);
        rc = SDDDST_ERR_SOLVER;
    }
done:
    if (handle) rocblas_destroy_handle(handle);
    for (void *p : {(void *)H, (void *)lam, (void *)E, (void *)obs, (void *)info, (void *)C, (void *)W, (void *)sp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace sdamd
