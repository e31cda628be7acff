// Restarted flexible GMRES in FP64 for A delta = g, with A in its CSR view
// (AxRow on the CSC index arrays: the pattern is symmetric) and a right
// preconditioner applied per direction (Jacobi by default, a sparse
// approximate inverse when one is installed).  Acceptance: the true relative
// residual ||g - A delta|| / ||g|| <= rel_tol, checked with an explicit SpMV
// at the end of every cycle; non-convergence returns SDDDST_ERR_SOLVER and
// the dispatcher falls back to the direct solver.
//
// Arnoldi uses classical Gram-Schmidt applied twice (CGS2): one kernel
// computes all j+1 projections at once, so each iteration costs one SpMV,
// two projection passes and one 8*(j+2)-byte device-to-host copy of the
// Hessenberg column; the Givens rotations run on the host.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <vector>

#include "context.h"
#include "precond_core.h"
#include "solvers.h"

namespace sdamd {

namespace {

constexpr int kRed = 64;    // workgroups of the reductions
constexpr int kMaxV = 64;   // most vectors one dots launch handles (restart + 1)

// y = A x, one wavefront per row.
__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                              const double *__restrict__ Av, const double *__restrict__ x,
                                              double *__restrict__ y)
{
    const int lane = (int)(threadIdx.x & 63);
    const int r = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (r >= n) return;
    double s = 0.0;
    for (int k = Ap[r] + lane; k < Ap[r + 1]; k += 64) s += Av[k] * x[Ai[k]];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[r] = s;
}

// out[v] (+)= sum_i V[v][i] * w[i] for v < nv in one launch: every block
// reduces its rows per v (wave shuffles, then LDS across the 4 waves) into
// partial[block][v]; the last block to finish (atomic ticket) adds the
// partials in block order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_dots(int n, int nv, const double *__restrict__ V,
                                              const double *__restrict__ w, double *__restrict__ partial,
                                              unsigned *__restrict__ ticket, double *__restrict__ out, int accumulate)
{
    __shared__ double wsum[4][kMaxV];
    __shared__ bool last;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int stride = (int)(gridDim.x * blockDim.x);
    const int i0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    for (int v = 0; v < nv; ++v) {
        double acc = 0.0;
        for (int i = i0; i < n; i += stride) acc += V[(size_t)v * n + i] * w[i];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) wsum[wave][v] = acc;
    }
    __syncthreads();
    for (int v = (int)threadIdx.x; v < nv; v += 256)
        partial[(size_t)blockIdx.x * nv + v] = ((wsum[0][v] + wsum[1][v]) + wsum[2][v]) + wsum[3][v];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int v = (int)threadIdx.x; v < nv; v += 256) {
        double sum = 0.0;
        for (int b = 0; b < (int)gridDim.x; ++b) sum += __builtin_nontemporal_load(&partial[(size_t)b * nv + v]);
        out[v] = accumulate ? out[v] + sum : sum;
    }
    if (threadIdx.x == 0) *ticket = 0u;
}

// The same reduction by one workgroup of 1024 lanes (systems up to 2^18
// rows): strided loads, wave sums, then the 16 wave partials of each v added
// in wave order; no atomics, one launch.  out[v] = V[v] . w; with_self also
// out[nv] = w . w; acc (nullable) receives (accumulate ? acc : 0) + out.
__global__ __launch_bounds__(1024) void k_dots1(int n, int nv, const double *__restrict__ V,
                                                const double *__restrict__ w, double *__restrict__ out,
                                                double *__restrict__ acc_out, int accumulate, int with_self)
{
    __shared__ double wsum[16][kMaxV];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    // four vectors x eight elements of loads in flight before their products;
    // w . w rides along with the first group
    double wsq = 0.0;
    for (int v0 = 0; v0 < nv; v0 += 4) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i0 = (int)threadIdx.x; i0 < n; i0 += 8 * 1024) {
            double wv[8], p[4][8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * 1024;
                wv[u] = i < n ? w[i] : 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) p[q][u] = (i < n && v0 + q < nv) ? V[(size_t)(v0 + q) * n + i] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int u = 0; u < 8; ++u) acc[q] += p[q][u] * wv[u];
            if (v0 == 0 && with_self)
#pragma unroll
                for (int u = 0; u < 8; ++u) wsq += wv[u] * wv[u];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double a = acc[q];
            for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
            if (lane == 0 && v0 + q < nv) wsum[wave][v0 + q] = a;
        }
    }
    const int nvt = nv + (with_self ? 1 : 0);
    if (with_self) {
        for (int off = 32; off > 0; off >>= 1) wsq += __shfl_down(wsq, off, 64);
        if (lane == 0) wsum[wave][nv] = wsq;
    }
    __syncthreads();
    for (int v = (int)threadIdx.x; v < nvt; v += 1024) {
        double sum = 0.0;
        for (int k = 0; k < 16; ++k) sum += wsum[k][v];
        out[v] = sum;
        if (acc_out && v < nv) acc_out[v] = accumulate ? acc_out[v] + sum : sum;
    }
}

// w -= sum_v h[v] V[v]   (h from the latest projection pass in hp)
__global__ void k_orth(int n, int nv, const double *__restrict__ V, const double *__restrict__ hp,
                       double *__restrict__ w)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = w[i];
    for (int v = 0; v < nv; ++v) acc -= hp[v] * V[(size_t)v * n + i];
    w[i] = acc;
}

__global__ void k_scale_copy(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) y[i] = a * x[i];
}

__global__ void k_jacobi_apply(int n, const double *__restrict__ dinv, const double *__restrict__ v,
                               double *__restrict__ z)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) z[i] = dinv[i] * v[i];
}

__global__ void k_jacobi_setup(int n, const int *__restrict__ diag, const double *__restrict__ Ax,
                               double *__restrict__ dinv)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) dinv[i] = 1.0 / Ax[diag[i]];
}

// Pair-block Jacobi (row bodies in precond_core.h), one wavefront per row.
__global__ __launch_bounds__(256) void k_pair_partner(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                      const int *__restrict__ diag, const double *__restrict__ Ax,
                                                      const double *__restrict__ AxRow, int *__restrict__ partner,
                                                      double *__restrict__ strength)
{
    const int i = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    pairblk::partner_row(i, Ap, Ai, diag, Ax, AxRow, partner, strength);
}

__global__ __launch_bounds__(256) void k_pair_blocks(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                     const int *__restrict__ diag, const double *__restrict__ Ax,
                                                     const double *__restrict__ AxRow, const int *__restrict__ partner,
                                                     const double *__restrict__ strength, double threshold,
                                                     double *__restrict__ binv, int *__restrict__ bcol)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    pairblk::block_row(i, Ap, Ai, diag, Ax, AxRow, partner, strength, threshold, binv, bcol);
}

__global__ void k_pair_apply(int n, const double *__restrict__ binv, const int *__restrict__ bcol,
                             const double *__restrict__ v, double *__restrict__ z)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int j = bcol[i];
    double r = binv[2 * i] * v[i];
    if (j >= 0) r += binv[2 * i + 1] * v[j];
    z[i] = r;
}

// hsum[v] = (accumulate ? hsum[v] : 0) + hp[v]
__global__ void k_add(int nv, const double *__restrict__ hp, double *__restrict__ hsum, int accumulate)
{
    for (int v = (int)threadIdx.x; v < nv; v += 64) hsum[v] = accumulate ? hsum[v] + hp[v] : hp[v];
}

// x += sum_j y[j] Z[j]
__global__ void k_combine(int n, int k, const double *__restrict__ Z, const double *__restrict__ yv,
                          double *__restrict__ x)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = x[i];
    for (int j = 0; j < k; ++j) acc += yv[j] * Z[(size_t)j * n + i];
    x[i] = acc;
}

// r = g - r (r holds A x on entry)
__global__ void k_resid(int n, const double *__restrict__ g, double *__restrict__ r)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) r[i] = g[i] - r[i];
}

}  // namespace

static inline int nb256(int n) { return (n + 255) / 256; }

// out[v] = V[v] . w for v < nv (and out[nv] = w . w with with_self);
// acc (nullable) += or = out[0..nv)
static void launch_dots(Context &c, Fgmres &f, int nv, const double *V, const double *w, double *out,
                        double *acc = nullptr, int accumulate = 0, int with_self = 0)
{
    if (f.n <= (1 << 18)) {
        hipLaunchKernelGGL(k_dots1, dim3(1), dim3(1024), 0, c.stream, f.n, nv, V, w, out, acc, accumulate,
                           with_self);
        return;
    }
    hipLaunchKernelGGL(k_dots, dim3(kRed), dim3(256), 0, c.stream, f.n, nv, V, w, f.red, f.ticket, out, 0);
    if (acc) hipLaunchKernelGGL(k_add, dim3(1), dim3(64), 0, c.stream, nv, out, acc, accumulate);
    if (with_self)
        hipLaunchKernelGGL(k_dots, dim3(kRed), dim3(256), 0, c.stream, f.n, 1, w, w, f.red, f.ticket, out + nv, 0);
}

int Fgmres::init(Context &c)
{
    if (V && n == c.n && m == c.solver_restart) return SDDDST_OK;
    release();
    n = c.n;
    m = c.solver_restart;
    SD_HIP(hipMalloc(&V, sizeof(double) * (size_t)(m + 1) * n));
    SD_HIP(hipMalloc(&Z, sizeof(double) * (size_t)m * n));
    SD_HIP(hipMalloc(&w, sizeof(double) * n));
    SD_HIP(hipMalloc(&dinv, sizeof(double) * n));
    SD_HIP(hipMalloc(&binv, sizeof(double) * 2 * n));
    SD_HIP(hipMalloc(&strength, sizeof(double) * n));
    SD_HIP(hipMalloc(&partner, sizeof(int) * n));
    SD_HIP(hipMalloc(&bcol, sizeof(int) * n));
    SD_CHECK(m + 1 <= kMaxV, SDDDST_ERR_ARG, "FGMRES restart %d exceeds %d", m, kMaxV - 1);
    SD_HIP(hipMalloc(&red, sizeof(double) * (size_t)kRed * (m + 2)));
    SD_HIP(hipMalloc(&ticket, sizeof(unsigned)));
    SD_HIP(hipMemsetAsync(ticket, 0, sizeof(unsigned), c.stream));
    SD_HIP(hipMalloc(&hcol, sizeof(double) * 2 * (m + 2)));
    SD_HIP(hipHostMalloc(&H, sizeof(double) * 2 * (m + 2), hipHostMallocDefault));
    return SDDDST_OK;
}

void Fgmres::release()
{
    if (V) (void)hipFree(V);
    if (Z) (void)hipFree(Z);
    if (w) (void)hipFree(w);
    if (dinv) (void)hipFree(dinv);
    if (red) (void)hipFree(red);
    if (hcol) (void)hipFree(hcol);
    if (H) (void)hipHostFree(H);
    if (ticket) (void)hipFree(ticket);
    for (void *p : {(void *)binv, (void *)strength, (void *)partner, (void *)bcol})
        if (p) (void)hipFree(p);
    binv = strength = nullptr;
    partner = bcol = nullptr;
    V = Z = w = dinv = red = hcol = H = nullptr;
    ticket = nullptr;
}

int Fgmres::pair_factor(Context &c)
{
    int rc = init(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pair_partner, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.Ax,
                       c.AxRow, partner, strength);
    hipLaunchKernelGGL(k_pair_blocks, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.Ax, c.AxRow,
                       partner, strength, 1e-3, binv, bcol);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

int Fgmres::factor(Context &c)
{
    int rc = init(c);
    if (rc) return rc;
    if (c.precond == SDDDST_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_jacobi_setup, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.diag, c.Ax, dinv);
        SD_HIP(hipGetLastError());
        return SDDDST_OK;
    }
    if ((rc = pair_factor(c))) return rc;
    if (c.precond == SDDDST_PRECOND_LEARNED) return spai_factor(c, *this);
    return SDDDST_OK;
}

int Fgmres::apply(Context &c, const double *v, double *z)
{
    const int precond = c.precond_override >= 0 ? c.precond_override : c.precond;
    if (precond == SDDDST_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_jacobi_apply, dim3(nb256(n)), dim3(256), 0, c.stream, n, dinv, v, z);
    else if (precond == SDDDST_PRECOND_LEARNED)
        return spai_apply(c, *this, v, z);
    else
        hipLaunchKernelGGL(k_pair_apply, dim3(nb256(n)), dim3(256), 0, c.stream, n, binv, bcol, v, z);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// CGS2: two passes of (h = V^T w, w -= V h) accumulated into hsum, then
// hsum[nv] = ||w||; one copy of the column to the host.
// The second pass also returns ||w||^2 of its input, so the new norm is
// ||w - V h2||^2 = ||w||^2 - ||h2||^2 (V orthonormal); h2 is tiny after the
// first pass, and when it is not (||h2||^2 > ||w||^2 / 16) the norm is
// computed explicitly.
static int arnoldi_orth(Context &c, Fgmres &f, int nv, double *hhost)
{
    const int n = f.n;
    double *hp = f.hcol;                // this pass's projections (+ w.w)
    double *hsum = f.hcol + (f.m + 2);  // accumulated
    for (int pass = 0; pass < 2; ++pass) {
        launch_dots(c, f, nv, f.V, f.w, hp, hsum, pass, pass);
        hipLaunchKernelGGL(k_orth, dim3(nb256(n)), dim3(256), 0, c.stream, n, nv, f.V, hp, f.w);
        SD_HIP(hipGetLastError());
    }
    // one copy of hcol: h2 and w.w at [0, nv], h at [m+2, m+2+nv)
    double *h2 = hhost + (f.m + 2);
    SD_HIP(hipMemcpyAsync(h2, hp, sizeof(double) * (f.m + 2 + nv), hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    for (int v = 0; v < nv; ++v) hhost[v] = h2[f.m + 2 + v];
    double s2 = 0.0;
    for (int v = 0; v < nv; ++v) s2 += h2[v] * h2[v];
    const double ww = h2[nv];
    if (s2 <= ww * 0.0625) {
        hhost[nv] = sqrt(ww - s2);
        return SDDDST_OK;
    }
    launch_dots(c, f, 1, f.w, f.w, hsum + nv);
    SD_HIP(hipGetLastError());
    SD_HIP(hipMemcpyAsync(hhost + nv, hsum + nv, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    hhost[nv] = sqrt(hhost[nv]);
    return SDDDST_OK;
}

static int norm2(Context &c, Fgmres &f, const double *v, double *out)
{
    launch_dots(c, f, 1, v, v, f.hcol);
    SD_HIP(hipGetLastError());
    SD_HIP(hipMemcpyAsync(f.H, f.hcol, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    *out = sqrt(f.H[0]);
    return SDDDST_OK;
}

int Fgmres::solve(Context &c)
{
    const int N = n;
    const double tol = c.solver_tol;
    // a learned-preconditioned solve gets one restart cycle; the caller
    // retries with the pair blocks when it does not converge in it
    const int precond = c.precond_override >= 0 ? c.precond_override : c.precond;
    const int max_it = precond == SDDDST_PRECOND_LEARNED ? std::min(c.solver_max_iter, m) : c.solver_max_iter;
    int rc;
    // delta = 0, r = g
    SD_HIP(hipMemsetAsync(c.delta, 0, sizeof(double) * N, c.stream));
    double gnorm;
    rc = norm2(c, *this, c.g, &gnorm);
    if (rc) return rc;
    c.last_iters = 0;
    if (gnorm == 0.0) {
        c.last_rel_res = 0.0;
        return SDDDST_OK;
    }
    std::vector<double> Hm((size_t)(m + 1) * m), cs(m), sn(m), s(m + 1), y(m), hh(3 * (m + 2) + 1);
    int total_it = 0;
    double rel = 1.0;
    const double *r0 = c.g;
    for (int cycle = 0; total_it < max_it; ++cycle) {
        // r = g - A delta
        if (cycle > 0) {
            hipLaunchKernelGGL(k_spmv, dim3((N + 3) / 4), dim3(256), 0, c.stream, N, c.Ap, c.Ai, c.AxRow, c.delta, w);
            hipLaunchKernelGGL(k_resid, dim3(nb256(N)), dim3(256), 0, c.stream, N, c.g, w);
            SD_HIP(hipGetLastError());
            r0 = w;
        }
        double beta;
        rc = norm2(c, *this, r0, &beta);
        if (rc) return rc;
        rel = beta / gnorm;
        if (rel <= tol) break;
        hipLaunchKernelGGL(k_scale_copy, dim3(nb256(N)), dim3(256), 0, c.stream, N, 1.0 / beta, r0, V);
        SD_HIP(hipGetLastError());
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        int k = 0;
        for (; k < m && total_it < max_it; ++k, ++total_it) {
            // z_k = M^-1 v_k ; w = A z_k
            rc = apply(c, V + (size_t)k * N, Z + (size_t)k * N);
            if (rc) return rc;
            hipLaunchKernelGGL(k_spmv, dim3((N + 3) / 4), dim3(256), 0, c.stream, N, c.Ap, c.Ai, c.AxRow,
                               Z + (size_t)k * N, w);
            SD_HIP(hipGetLastError());
            rc = arnoldi_orth(c, *this, k + 1, hh.data());
            if (rc) return rc;
            const double hnext = hh[k + 1];
            if (hnext > 0)
                hipLaunchKernelGGL(k_scale_copy, dim3(nb256(N)), dim3(256), 0, c.stream, N, 1.0 / hnext, w,
                                   V + (size_t)(k + 1) * N);
            // apply previous rotations to the new column
            for (int i = 0; i < k; ++i) {
                const double t = cs[i] * hh[i] + sn[i] * hh[i + 1];
                hh[i + 1] = -sn[i] * hh[i] + cs[i] * hh[i + 1];
                hh[i] = t;
            }
            const double denom = hypot(hh[k], hh[k + 1]);
            cs[k] = denom > 0 ? hh[k] / denom : 1.0;
            sn[k] = denom > 0 ? hh[k + 1] / denom : 0.0;
            hh[k] = denom;
            hh[k + 1] = 0.0;
            s[k + 1] = -sn[k] * s[k];
            s[k] = cs[k] * s[k];
            for (int i = 0; i <= k; ++i) Hm[(size_t)i * m + k] = hh[i];
            if (fabs(s[k + 1]) / gnorm <= 0.5 * tol || hnext == 0.0) {
                ++k;
                ++total_it;
                break;
            }
        }
        // back substitution H y = s
        for (int i = k - 1; i >= 0; --i) {
            double acc = s[i];
            for (int j = i + 1; j < k; ++j) acc -= Hm[(size_t)i * m + j] * y[j];
            y[i] = acc / Hm[(size_t)i * m + i];
        }
        SD_HIP(hipMemcpyAsync(hcol, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_combine, dim3(nb256(N)), dim3(256), 0, c.stream, N, k, Z, hcol, c.delta);
        SD_HIP(hipGetLastError());
        SD_HIP(hipStreamSynchronize(c.stream));
    }
    // true residual
    hipLaunchKernelGGL(k_spmv, dim3((N + 3) / 4), dim3(256), 0, c.stream, N, c.Ap, c.Ai, c.AxRow, c.delta, w);
    hipLaunchKernelGGL(k_resid, dim3(nb256(N)), dim3(256), 0, c.stream, N, c.g, w);
    SD_HIP(hipGetLastError());
    double rn;
    rc = norm2(c, *this, w, &rn);
    if (rc) return rc;
    c.last_iters = total_it;
    c.last_rel_res = rn / gnorm;
    if (!(c.last_rel_res <= tol)) {
        set_error("FGMRES: relative residual %.3e after %d iterations", c.last_rel_res, total_it);
        return SDDDST_ERR_SOLVER;
    }
    return SDDDST_OK;
}

}  // namespace sdamd
