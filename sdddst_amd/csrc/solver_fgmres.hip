// Restarted flexible GMRES in FP64 for A delta = g, with A in its CSR view
// (AxRow on the CSC index arrays: the pattern is symmetric) and a right
// preconditioner applied per direction (Jacobi, pair blocks, or the learned
// sparse approximate inverse).  Acceptance: the true relative residual
// ||g - A delta|| / ||g|| <= rel_tol, checked with an explicit SpMV at the
// end of the solve; non-convergence returns SDDDST_ERR_SOLVER and the
// dispatcher falls back to the direct solver.
//
// Arnoldi uses classical Gram-Schmidt applied twice (CGS2), each pass one
// kernel over all j+1 basis vectors.  An iteration costs the SpMV, two
// projection launches and one device-to-host copy of the Hessenberg column;
// the Givens rotations run on the host.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <vector>

#include "context.h"
#include "precond_core.h"
#include "solvers.h"

namespace sdamd {

namespace {

constexpr int kRedMax = 256;  // most workgroups of one projection launch
constexpr int kMaxV = 64;   // most vectors one dots launch handles (restart + 1)

// One row's dot product by one wavefront: lane l adds the products of
// entries l, l + 64, ... in order, and the lanes are combined by a fixed
// butterfly.  Each lane issues the value and index loads of eight of its
// entries before their products, so a row costs one or two memory round
// trips instead of one per 64 entries; the order of the additions is that
// of the plain loop.
constexpr int kRowUnroll = 8;
template <typename XRead>
__device__ __forceinline__ double row_dot(int b, int e, int lane, const int *__restrict__ Ai,
                                          const double *__restrict__ Av, XRead xr)
{
    double s = 0.0;
    for (int k0 = b + lane; k0 < e; k0 += 64 * kRowUnroll) {
        double a[kRowUnroll];
        int c[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u) {
            const int k = k0 + 64 * u;
            a[u] = k < e ? Av[k] : 0.0;
            c[u] = k < e ? Ai[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k0 + 64 * u < e) s += a[u] * xr(c[u]);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    return s;
}

// y = A x, one wavefront per row.
__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                              const double *__restrict__ Av, const double *__restrict__ x,
                                              double *__restrict__ y)
{
    const int lane = (int)(threadIdx.x & 63);
    const int r = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (r >= n) return;
    const double s = row_dot(Ap[r], Ap[r + 1], lane, Ai, Av, [&](int i) { return x[i]; });
    if (lane == 0) y[r] = s;
}

// The same product with x staged in LDS (systems up to kSpmvLdsMax rows):
// one persistent 1024-lane workgroup per CU copies x once, then its
// wavefronts take rows with stride 16 * grid; the gathers of x hit LDS
// instead of the cache hierarchy.  Each row's sum is formed exactly as in
// k_spmv.
constexpr int kSpmvLdsMax = 20480;   // 160 KiB of FP64
__global__ __launch_bounds__(1024) void k_spmv_lds(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                   const double *__restrict__ Av, const double *__restrict__ x,
                                                   double *__restrict__ y)
{
    extern __shared__ double xs[];
    for (int i = (int)threadIdx.x; i < n; i += 1024) xs[i] = x[i];
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63);
    for (int r = (int)(blockIdx.x * 16 + (threadIdx.x >> 6)); r < n; r += (int)gridDim.x * 16) {
        const double s = row_dot(Ap[r], Ap[r + 1], lane, Ai, Av, [&](int i) { return xs[i]; });
        if (lane == 0) y[r] = s;
    }
}

__device__ __forceinline__ double wave_sum(double s)
{
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// One classical Gram-Schmidt pass of the Arnoldi step in one launch.  With
// hsub (nullable) w first becomes w - sum_v hsub[v] V[v] (stored back, the
// subtraction of k_orth); then out[v] = V[v] . w for v < nv and, with
// with_self, out[nv] = w . w; acc (nullable) receives
// (accumulate ? acc : 0) + out[0..nv); host_out (nullable, mapped pinned
// memory) receives out at [0, nvt) and acc at [host_acc_off, +nv).  Each
// workgroup reduces its rows (wave
// butterflies, the 4 waves in order) into partial[block][0..nv]; the last
// workgroup to arrive (atomic ticket) adds the partials in a fixed tree over
// the block index, so the sums do not depend on scheduling.
__global__ __launch_bounds__(256) void k_proj(int n, int nv, const double *__restrict__ V, double *__restrict__ w,
                                              const double *__restrict__ hsub, double *__restrict__ partial,
                                              int pstride, unsigned *__restrict__ ticket, double *__restrict__ out,
                                              double *__restrict__ acc_out, int accumulate, int with_self,
                                              double *__restrict__ host_out, int host_acc_off)
{
    __shared__ double wsum[4][kMaxV + 1];
    __shared__ bool last;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int stride = (int)(gridDim.x * 256);
    const int i0 = (int)(blockIdx.x * 256 + threadIdx.x);
    if (hsub) {
        for (int i = i0; i < n; i += stride) {
            double a = w[i];
            for (int v = 0; v < nv; ++v) a -= hsub[v] * V[(size_t)v * n + i];
            w[i] = a;
        }
    }
    const int nvt = nv + (with_self ? 1 : 0);
    // four projections per sweep: their loads are in flight together
    for (int v0 = 0; v0 < nvt; v0 += 4) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = i0; i < n; i += stride) {
            const double wi = w[i];
            double p[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int v = v0 + q;
                p[q] = v < nv ? V[(size_t)v * n + i] : (v == nv ? wi : 0.0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += p[q] * wi;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double t = wave_sum(acc[q]);
            if (lane == 0 && v0 + q < nvt) wsum[wave][v0 + q] = t;
        }
    }
    __syncthreads();
    for (int v = (int)threadIdx.x; v < nvt; v += 256)
        partial[(size_t)blockIdx.x * pstride + v] = ((wsum[0][v] + wsum[1][v]) + wsum[2][v]) + wsum[3][v];
    // the barrier completes the workgroup's stores; one lane releases them
    // to the device before taking the ticket, the last workgroup acquires
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int v = wave; v < nvt; v += 4) {
        double t = 0.0;
        for (int b = lane; b < (int)gridDim.x; b += 64)
            t += __builtin_nontemporal_load(&partial[(size_t)b * pstride + v]);
        t = wave_sum(t);
        if (lane == 0) {
            out[v] = t;
            const double a = (acc_out && v < nv) ? (accumulate ? acc_out[v] + t : t) : 0.0;
            if (acc_out && v < nv) acc_out[v] = a;
            // the host's copy, in mapped pinned memory (read after the host's
            // completion wait)
            if (host_out) {
                host_out[v] = t;
                if (acc_out && v < nv) host_out[host_acc_off + v] = a;
            }
        }
    }
    if (threadIdx.x == 0) *ticket = 0u;
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

// y = a * (w - sum_v h[v] V[v]): the second Gram-Schmidt subtraction and the
// normalisation of the next basis vector in one pass (w is left unchanged).
__global__ void k_orth_scale(int n, int nv, const double *__restrict__ V, const double *__restrict__ hp,
                             const double *__restrict__ w, double a, double *__restrict__ y)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = w[i];
    for (int v = 0; v < nv; ++v) acc -= hp[v] * V[(size_t)v * n + i];
    y[i] = a * acc;
}

__global__ void k_scale_copy(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) y[i] = a * x[i];
}

// y = x / sqrt(*nrm2) (0 when the norm is 0), the norm read on the device
__global__ void k_scale_by_norm(int n, const double *__restrict__ nrm2, const double *__restrict__ x,
                                double *__restrict__ y)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double b = sqrt(*nrm2);
    y[i] = (b > 0 ? 1.0 / b : 0.0) * x[i];
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

// x += sum_j y[j] Z[j]
struct Coefs {
    double v[kMaxV];
};

// x += sum_j y[j] Z[j], y passed by value
__global__ void k_combine(int n, int k, const double *__restrict__ Z, Coefs yv, double *__restrict__ x)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = x[i];
    for (int j = 0; j < k; ++j) acc += yv.v[j] * Z[(size_t)j * n + i];
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

static int spmv_grid()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, cu = 0;
        SD_HIP(hipGetDevice(&dev));
        SD_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        SD_HIP(hipFuncSetAttribute((const void *)k_spmv_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * kSpmvLdsMax)));
        cus = cu < 1 ? 1 : (cu > kRedMax ? kRedMax : cu);
    }
    return cus;
}

// y = A x on the CSR view (AxRow): LDS-staged x when it fits.
static int launch_spmv(Context &c, const double *x, double *y)
{
    const int n = c.n;
    if (n <= kSpmvLdsMax) {
        const int cus = spmv_grid();
        if (cus < 0) return cus;
        hipLaunchKernelGGL(k_spmv_lds, dim3(cus), dim3(1024), sizeof(double) * n, c.stream, n, c.Ap, c.Ai, c.AxRow, x,
                           y);
    } else {
        hipLaunchKernelGGL(k_spmv, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.AxRow, x, y);
    }
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

static inline int proj_blocks(int n)
{
    const int b = nb256(n);
    return b < kRedMax ? (b > 0 ? b : 1) : kRedMax;
}

// One projection pass (k_proj): with hsub, w -= V hsub first; out[v] = V[v].w
// for v < nv (and out[nv] = w.w with with_self); acc (nullable) += or = out.
static void launch_proj(Context &c, Fgmres &f, int nv, const double *V, double *w, const double *hsub,
                        double *out, double *acc = nullptr, int accumulate = 0, int with_self = 0,
                        double *host_out = nullptr, int host_acc_off = 0)
{
    hipLaunchKernelGGL(k_proj, dim3(proj_blocks(f.n)), dim3(256), 0, c.stream, f.n, nv, V, w, hsub, f.red, f.m + 2,
                       f.ticket, out, acc, accumulate, with_self, host_out, host_acc_off);
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
    SD_HIP(hipMalloc(&bcol, sizeof(int) * n));
    SD_CHECK(m + 1 <= kMaxV, SDDDST_ERR_ARG, "FGMRES restart %d exceeds %d", m, kMaxV - 1);
    SD_HIP(hipMalloc(&red, sizeof(double) * (size_t)kRedMax * (m + 2)));
    SD_HIP(hipMalloc(&ticket, sizeof(unsigned)));
    SD_HIP(hipMemsetAsync(ticket, 0, sizeof(unsigned), c.stream));
    SD_HIP(hipMalloc(&hcol, sizeof(double) * (3 * (m + 2) + 2)));
    // results the host reads after a synchronisation: written by the
    // reductions' last workgroup into mapped, coherent pinned memory
    SD_HIP(hipHostMalloc(&H, sizeof(double) * (3 * (m + 2) + 2), hipHostMallocMapped | hipHostMallocCoherent));
    SD_HIP(hipHostGetDevicePointer((void **)&Hdev, H, 0));
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
    for (void *p : {(void *)binv, (void *)bcol})
        if (p) (void)hipFree(p);
    binv = nullptr;
    bcol = nullptr;
    V = Z = w = dinv = red = hcol = H = Hdev = nullptr;
    ticket = nullptr;
}

int Fgmres::pair_factor(Context &c)
{
    int rc = init(c);
    if (rc) return rc;
    // the fused scaling pass of the assembly leaves the partners of this
    // matrix in the context; otherwise (compacted pattern) find them here
    if (!c.partner_valid) {
        hipLaunchKernelGGL(k_pair_partner, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.Ax,
                           c.AxRow, c.partner, c.strength);
        c.partner_valid = true;
    }
    hipLaunchKernelGGL(k_pair_blocks, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.Ax, c.AxRow,
                       c.partner, c.strength, 1e-3, binv, bcol);
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
// Arnoldi step for the new vector w against V[0..nv): CGS2 (two projection
// passes, the first subtraction fused into the second pass).  On return
// hhost[0..nv) = h = h1 + h2 and hhost[nv] = ||w - V h||, from
// ||w'||^2 - ||h2||^2 (w' = w - V h1, orthogonal to V up to h2) unless the
// second pass removed more than a quarter of ||w'||, when the norm is taken
// explicitly.  next (nullable) receives (w - V h) / hhost[nv] when that is
// positive; w then holds w' (or w - V h after the explicit norm).
static int arnoldi_orth(Context &c, Fgmres &f, int nv, double *hhost, double *next)
{
    const int n = f.n;
    double *h1 = f.hcol;                      // first pass
    double *h2 = f.hcol + (f.m + 2);          // second pass (+ w'.w')
    double *hsum = f.hcol + 2 * (f.m + 2);    // h1 + h2
    launch_proj(c, f, nv, f.V, f.w, nullptr, h1, hsum, 0, 0);
    // the second pass also writes h2, w'.w' (at [0, nv]) and h (at
    // [m+2, m+2+nv)) into the mapped pinned buffer
    launch_proj(c, f, nv, f.V, f.w, h1, h2, hsum, 1, 1, f.Hdev, f.m + 2);
    SD_HIP(hipGetLastError());
    int rc = host_wait(c);
    if (rc) return rc;
    const double *hc = f.H;
    for (int v = 0; v < nv; ++v) hhost[v] = hc[f.m + 2 + v];
    double s2 = 0.0;
    for (int v = 0; v < nv; ++v) s2 += hc[v] * hc[v];
    const double ww = hc[nv];
    if (s2 <= ww * 0.0625) {
        const double hn = sqrt(ww - s2);
        hhost[nv] = hn;
        if (next && hn > 0)
            hipLaunchKernelGGL(k_orth_scale, dim3(nb256(n)), dim3(256), 0, c.stream, n, nv, f.V, h2, f.w, 1.0 / hn,
                               next);
        SD_HIP(hipGetLastError());
        return SDDDST_OK;
    }
    hipLaunchKernelGGL(k_orth, dim3(nb256(n)), dim3(256), 0, c.stream, n, nv, f.V, h2, f.w);
    launch_proj(c, f, 0, f.V, f.w, nullptr, hsum + nv, nullptr, 0, 1, f.Hdev);
    SD_HIP(hipGetLastError());
    if ((rc = host_wait(c))) return rc;
    hhost[nv] = sqrt(f.H[0]);
    if (next && hhost[nv] > 0)
        hipLaunchKernelGGL(k_scale_copy, dim3(nb256(n)), dim3(256), 0, c.stream, n, 1.0 / hhost[nv], f.w, next);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

static int norm2(Context &c, Fgmres &f, const double *v, double *out)
{
    launch_proj(c, f, 0, f.V, const_cast<double *>(v), nullptr, f.hcol, nullptr, 0, 1, f.Hdev);
    SD_HIP(hipGetLastError());
    const int rc = host_wait(c);
    if (rc) return rc;
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
    // delta = 0, r = g; ||g||^2 lands in the device slot gs and its mapped
    // copy, and v_0 = g / ||g|| is formed on the device, so the host learns
    // ||g|| at the first iteration's synchronisation
    SD_HIP(hipMemsetAsync(c.delta, 0, sizeof(double) * N, c.stream));
    double *gs = hcol + 3 * (m + 2);
    launch_proj(c, *this, 0, V, const_cast<double *>(c.g), nullptr, gs, nullptr, 0, 1, Hdev + 3 * (m + 2));
    hipLaunchKernelGGL(k_scale_by_norm, dim3(nb256(N)), dim3(256), 0, c.stream, N, gs, c.g, V);
    SD_HIP(hipGetLastError());
    double gnorm = -1.0;
    c.last_iters = 0;
    std::vector<double> Hm((size_t)(m + 1) * m), cs(m), sn(m), s(m + 1), y(m), hh(3 * (m + 2) + 1);
    int total_it = 0;
    double rel = 1.0;
    bool checked = false;   // rel is the true residual of the current delta
    const double *r0 = c.g;
    for (int cycle = 0;; ++cycle) {
        double beta = gnorm;   // delta = 0 in cycle 0: r = g
        if (cycle > 0) {
            // r = g - A delta, the true residual of the last cycle's delta
            if ((rc = launch_spmv(c, c.delta, w))) return rc;
            hipLaunchKernelGGL(k_resid, dim3(nb256(N)), dim3(256), 0, c.stream, N, c.g, w);
            SD_HIP(hipGetLastError());
            r0 = w;
            rc = norm2(c, *this, r0, &beta);
            if (rc) return rc;
            rel = beta / gnorm;
            checked = true;
            if (rel <= tol) break;
        }
        if (total_it >= max_it) break;
        checked = false;
        if (cycle > 0) {
            hipLaunchKernelGGL(k_scale_copy, dim3(nb256(N)), dim3(256), 0, c.stream, N, 1.0 / beta, r0, V);
            SD_HIP(hipGetLastError());
        }
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        int k = 0;
        for (; k < m && total_it < max_it; ++k, ++total_it) {
            // z_k = M^-1 v_k ; w = A z_k
            rc = apply(c, V + (size_t)k * N, Z + (size_t)k * N);
            if (rc) return rc;
            if ((rc = launch_spmv(c, Z + (size_t)k * N, w))) return rc;
            rc = arnoldi_orth(c, *this, k + 1, hh.data(), V + (size_t)(k + 1) * N);
            if (rc) return rc;
            if (gnorm < 0) {
                // first synchronisation of the solve: ||g|| is on the host
                gnorm = sqrt(H[3 * (m + 2)]);
                if (gnorm == 0.0) {
                    c.last_rel_res = 0.0;
                    return SDDDST_OK;
                }
                s[0] = beta = gnorm;
            }
            const double hnext = hh[k + 1];
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
        Coefs yc{};
        for (int j = 0; j < k; ++j) yc.v[j] = y[j];
        hipLaunchKernelGGL(k_combine, dim3(nb256(N)), dim3(256), 0, c.stream, N, k, Z, yc, c.delta);
        SD_HIP(hipGetLastError());
    }
    // the true residual, unless the last check already measured this delta
    if (!checked) {
        if ((rc = launch_spmv(c, c.delta, w))) return rc;
        hipLaunchKernelGGL(k_resid, dim3(nb256(N)), dim3(256), 0, c.stream, N, c.g, w);
        SD_HIP(hipGetLastError());
        double rn;
        rc = norm2(c, *this, w, &rn);
        if (rc) return rc;
        rel = rn / gnorm;
    }
    c.last_iters = total_it;
    c.last_rel_res = rel;
    if (!(c.last_rel_res <= tol)) {
        set_error("FGMRES: relative residual %.3e after %d iterations", c.last_rel_res, total_it);
        return SDDDST_ERR_SOLVER;
    }
    return SDDDST_OK;
}

}  // namespace sdamd
