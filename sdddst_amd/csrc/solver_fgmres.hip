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

// Stop codes of a device-driven restart cycle.  The kernels of an iteration
// run while the cycle runs; the closing kernels (solution update and true
// residual) once it has converged, broken down or used its iteration budget.
constexpr int kRunning = 0, kConverged = 1, kBreakdown = 2, kNeedNorm = 3, kZeroRhs = 4, kBudget = 5;
constexpr int kGateIter = 0, kGateClose = 1;

__device__ __forceinline__ bool gated(const int *gate, int mode)
{
    if (!gate) return false;
    const int s = *gate;
    return mode == kGateIter ? s != kRunning : !(s == kConverged || s == kBreakdown || s == kBudget);
}

}  // namespace

// The cycle's Givens state on the device (one lane updates it per iteration).
struct KrylovDev {
    double gnorm;             // ||g||
    double scale;             // 1 / h_{k+1,k} of the last step
    double s[kMaxV + 1];      // rotated right-hand side
    double cs[kMaxV], sn[kMaxV];
    double Hm[kMaxV * kMaxV]; // R of the rotated Hessenberg matrix, row-major
    int k, kmax, stop;
};

// What the host reads after a wait (res: the cycle's residual estimate
// relative to ||g||).
struct KrylovHost {
    double gnorm, res;
    int k, stop;
};

namespace {

// One row's dot product by one wavefront: lane l adds the products of
// entries l, l + 64, ... in order, and the lanes are combined by a fixed
// butterfly.  Each lane issues the value and index loads of eight of its
// entries before their products, so a row costs one or two memory round
// trips instead of one per 64 entries; the order of the additions is that
// of the plain loop.
constexpr int kRowUnroll = 8;
template <typename Index, typename XRead>
__device__ __forceinline__ double row_dot(int b, int e, int lane, const Index *__restrict__ Ai,
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
            c[u] = k < e ? (int)Ai[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k0 + 64 * u < e) s += a[u] * xr(c[u]);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    return s;
}

// y = A x, one wavefront per row.  Index: int, or unsigned short (the
// 16-bit copy of the row indices, N <= 65536).
template <typename Index>
__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ Ap, const Index *__restrict__ Ai,
                                              const double *__restrict__ Av, const double *__restrict__ x,
                                              double *__restrict__ y, const int *__restrict__ gate, int gmode,
                                              const double *__restrict__ gsub)
{
    if (gated(gate, gmode)) return;
    const int lane = (int)(threadIdx.x & 63);
    const int r = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (r >= n) return;
    const double s = row_dot(Ap[r], Ap[r + 1], lane, Ai, Av, [&](int i) { return x[i]; });
    if (lane == 0) y[r] = gsub ? gsub[r] - s : s;
}

// The same product with x staged in LDS (systems up to kSpmvLdsMax rows):
// one persistent 1024-lane workgroup per CU copies x once, then its
// wavefronts take rows with stride 16 * grid; the gathers of x hit LDS
// instead of the cache hierarchy.  Each row's sum is formed exactly as in
// k_spmv.  With nrm (the closing product of a restart cycle, y = g - A
// delta) the kernel also forms ||y||^2: each wavefront adds the squares of
// its rows in row order, the workgroup its 16 wavefronts in order into
// partial[block], and the last workgroup to take the ticket adds the
// partials in block order into *nrm (and nrm_host, mapped host memory) --
// deterministic, and one launch fewer per cycle than a separate reduction.
constexpr int kSpmvLdsMax = 20448;   // 160 KiB of FP64 with 32 words of reduction scratch
template <typename Index>
__global__ __launch_bounds__(1024) void k_spmv_lds(int n, const int *__restrict__ Ap, const Index *__restrict__ Ai,
                                                   const double *__restrict__ Av, const double *__restrict__ x,
                                                   double *__restrict__ y, const int *__restrict__ gate, int gmode,
                                                   const double *__restrict__ gsub, double *__restrict__ partial,
                                                   unsigned *__restrict__ ticket, double *__restrict__ nrm,
                                                   double *__restrict__ nrm_host)
{
    if (gated(gate, gmode)) return;   // every workgroup: the ticket stays untouched
    extern __shared__ double xs[];
    double *wsum = xs + n;            // 16 wavefront sums, then the last-workgroup flag
    for (int i = (int)threadIdx.x; i < n; i += 1024) xs[i] = x[i];
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    double ss = 0.0;
    for (int r = (int)(blockIdx.x * 16 + wave); r < n; r += (int)gridDim.x * 16) {
        const double s = row_dot(Ap[r], Ap[r + 1], lane, Ai, Av, [&](int i) { return xs[i]; });
        if (lane == 0) {
            const double v = gsub ? gsub[r] - s : s;
            y[r] = v;
            ss += v * v;
        }
    }
    if (!nrm) return;
    if (lane == 0) wsum[wave] = ss;
    __syncthreads();
    int *last = reinterpret_cast<int *>(wsum + 16);
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += wsum[w];
        partial[blockIdx.x] = t;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        *last = atomicAdd(ticket, 1u) == gridDim.x - 1 ? 1 : 0;
    }
    __syncthreads();
    if (!*last || wave != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double t = 0.0;
    for (int b = lane; b < (int)gridDim.x; b += 64) t += __builtin_nontemporal_load(&partial[b]);
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) {
        *nrm = t;
        if (nrm_host) *nrm_host = t;
        *ticket = 0u;
    }
}

__device__ __forceinline__ double wave_sum(double s)
{
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Iteration k of the cycle after the two projection passes (h2 with
// ||w'||^2 at h2[nv], h = h1 + h2 in hs): the new norm ||w'||^2 - ||h2||^2
// (or hn_given, taken explicitly by the host when the second pass removed
// more than a quarter of ||w'||), the previous rotations applied to the new
// column, the new rotation, the residual estimate and the stop decision.
// One lane; hh is LDS scratch of kMaxV + 1 doubles.
__device__ void arnoldi_update(KrylovDev *__restrict__ st, KrylovHost *__restrict__ kh, const double *h2,
                               const double *hs, int k, double halftol, double hn_given, double *hh)
{
    const int nv = k + 1;
    double hn = hn_given;
    if (hn_given < 0.0) {
        double s2 = 0.0;
        for (int v = 0; v < nv; ++v) s2 += h2[v] * h2[v];
        const double ww = h2[nv];
        if (s2 > ww * 0.0625) {
            st->stop = kNeedNorm;
            kh->k = k;
            kh->stop = kNeedNorm;
            __threadfence_system();
            return;
        }
        hn = sqrt(ww - s2);
    }
    for (int v = 0; v < nv; ++v) hh[v] = hs[v];
    hh[nv] = hn;
    for (int i = 0; i < k; ++i) {
        const double t = st->cs[i] * hh[i] + st->sn[i] * hh[i + 1];
        hh[i + 1] = -st->sn[i] * hh[i] + st->cs[i] * hh[i + 1];
        hh[i] = t;
    }
    const double denom = hypot(hh[k], hh[k + 1]);
    const double ck = denom > 0 ? hh[k] / denom : 1.0, sk = denom > 0 ? hh[k + 1] / denom : 0.0;
    st->cs[k] = ck;
    st->sn[k] = sk;
    hh[k] = denom;
    const double sk1 = -sk * st->s[k];
    st->s[k + 1] = sk1;
    st->s[k] = ck * st->s[k];
    for (int i = 0; i <= k; ++i) st->Hm[i * kMaxV + k] = hh[i];
    st->k = k + 1;
    st->scale = hn > 0 ? 1.0 / hn : 0.0;
    int stop = kRunning;
    if (fabs(sk1) / st->gnorm <= halftol || hn == 0.0) stop = hn == 0.0 ? kBreakdown : kConverged;
    else if (k + 1 >= st->kmax) stop = kBudget;
    st->stop = stop;
    kh->res = fabs(sk1) / st->gnorm;
    kh->k = k + 1;
    kh->stop = stop;
    __threadfence_system();
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
                                              double *__restrict__ host_out, int host_acc_off,
                                              const int *__restrict__ gate, int gmode, KrylovDev *__restrict__ st,
                                              KrylovHost *__restrict__ kh, double halftol)
{
    if (gated(gate, gmode)) return;   // every workgroup: the ticket stays untouched
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
    // the second pass of an Arnoldi step: the cycle's update from h2 (out,
    // with ||w'||^2 at out[nv]) and h (acc_out), by one lane of this workgroup
    if (st) {
        __syncthreads();
        if (threadIdx.x == 0) arnoldi_update(st, kh, out, acc_out, nv - 1, halftol, -1.0, &wsum[0][0]);
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

// (the diagonal is the same in the row- and column-major values)
__global__ void k_jacobi_setup(int n, const int *__restrict__ diag, const double *__restrict__ AxRow,
                               double *__restrict__ dinv)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) dinv[i] = 1.0 / AxRow[diag[i]];
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
// Start of a restart cycle: v_0 = r / beta with beta = sqrt(*beta2), and
// s = (beta, 0, ...); in cycle 0 (r = g) beta is also ||g||.
__global__ void k_cycle_start(int n, const double *__restrict__ r, const double *__restrict__ beta2,
                              double *__restrict__ v0, KrylovDev *__restrict__ st, KrylovHost *__restrict__ kh,
                              int first, int kmax)
{
    const double beta = sqrt(*beta2);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) v0[i] = (beta > 0 ? 1.0 / beta : 0.0) * r[i];
    if (i != 0) return;
    if (first) {
        st->gnorm = beta;
        kh->gnorm = beta;
    }
    st->s[0] = beta;
    st->k = 0;
    st->kmax = kmax;
    st->stop = st->gnorm == 0.0 ? kZeroRhs : kRunning;
    kh->res = st->gnorm > 0.0 ? beta / st->gnorm : 0.0;
    kh->k = 0;
    kh->stop = st->stop;
    __threadfence_system();
}

// The step resumed with an explicit norm (the host's path for kNeedNorm).
__global__ void k_arnoldi_step(KrylovDev *__restrict__ st, KrylovHost *__restrict__ kh, const double *__restrict__ hcol,
                               int m, int k, double halftol, double hn_given)
{
    __shared__ double hh[kMaxV + 1];
    if (threadIdx.x != 0) return;
    st->stop = kRunning;
    arnoldi_update(st, kh, hcol + (m + 2), hcol + 2 * (m + 2), k, halftol, hn_given, hh);
}

// y = scale * (w - sum_v h2[v] V[v]): the next basis vector, scale from the
// step; skipped once the cycle has stopped.
__global__ void k_orth_scale_dev(int n, int nv, const double *__restrict__ V, const double *__restrict__ hp,
                                 const double *__restrict__ w, const KrylovDev *__restrict__ st,
                                 double *__restrict__ y)
{
    if (st->stop != kRunning) return;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = w[i];
    for (int v = 0; v < nv; ++v) acc -= hp[v] * V[(size_t)v * n + i];
    y[i] = st->scale * acc;
}

// x += sum_j y[j] Z[j] once the cycle has ended, with R y = s solved by back
// substitution by one lane of each workgroup (k <= 63: a few hundred flops).
__global__ void k_combine_dev(int n, const double *__restrict__ Z, const KrylovDev *__restrict__ st,
                              double *__restrict__ x)
{
    __shared__ double y[kMaxV];
    if (gated(&st->stop, kGateClose)) return;
    const int k = st->k;
    if (threadIdx.x == 0) {
        for (int i = k - 1; i >= 0; --i) {
            double acc = st->s[i];
            for (int j = i + 1; j < k; ++j) acc -= st->Hm[i * kMaxV + j] * y[j];
            y[i] = acc / st->Hm[i * kMaxV + i];
        }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double acc = x[i];
    for (int j = 0; j < k; ++j) acc += y[j] * Z[(size_t)j * n + i];
    x[i] = acc;
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
        SD_HIP(hipFuncSetAttribute((const void *)k_spmv_lds<int>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * (kSpmvLdsMax + 32))));
        SD_HIP(hipFuncSetAttribute((const void *)k_spmv_lds<unsigned short>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * (kSpmvLdsMax + 32))));
        cus = cu < 1 ? 1 : (cu > kRedMax ? kRedMax : cu);
    }
    return cus;
}

// y = A x on the CSR view (AxRow): LDS-staged x when it fits.  With f and
// nrm: also ||y||^2 into nrm (and nrm_host) when the LDS kernel runs
// (*fused set), else the caller reduces.
static int launch_spmv(Context &c, const double *x, double *y, const int *gate = nullptr, int gmode = kGateIter,
                       const double *gsub = nullptr, Fgmres *f = nullptr, double *nrm = nullptr,
                       double *nrm_host = nullptr, bool *fused = nullptr)
{
    const int n = c.n;
    const bool i16 = c.ai16_valid;
    if (fused) *fused = false;
    if (n <= kSpmvLdsMax) {
        const int cus = spmv_grid();
        if (cus < 0) return cus;
        const bool red = f && nrm;
        if (fused) *fused = red;
        const size_t lds = sizeof(double) * (n + 32);
        double *part = red ? f->red : nullptr;
        unsigned *tk = red ? f->ticket : nullptr;
        double *nm = red ? nrm : nullptr, *nh = red ? nrm_host : nullptr;
        if (i16)
            hipLaunchKernelGGL(k_spmv_lds<unsigned short>, dim3(cus), dim3(1024), lds, c.stream, n, c.Ap, c.Ai16,
                               c.AxRow, x, y, gate, gmode, gsub, part, tk, nm, nh);
        else
            hipLaunchKernelGGL(k_spmv_lds<int>, dim3(cus), dim3(1024), lds, c.stream, n, c.Ap, c.Ai, c.AxRow, x, y,
                               gate, gmode, gsub, part, tk, nm, nh);
    } else if (i16) {
        hipLaunchKernelGGL(k_spmv<unsigned short>, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai16,
                           c.AxRow, x, y, gate, gmode, gsub);
    } else {
        hipLaunchKernelGGL(k_spmv<int>, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.AxRow, x, y,
                           gate, gmode, gsub);
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
                        double *host_out = nullptr, int host_acc_off = 0, const int *gate = nullptr,
                        int gmode = kGateIter, double halftol = -1.0)
{
    // halftol >= 0: the second pass of an Arnoldi step, which also updates the cycle
    hipLaunchKernelGGL(k_proj, dim3(proj_blocks(f.n)), dim3(256), 0, c.stream, f.n, nv, V, w, hsub, f.red, f.m + 2,
                       f.ticket, out, acc, accumulate, with_self, host_out, host_acc_off, gate, gmode,
                       halftol >= 0.0 ? f.kst : nullptr, f.khdev, halftol);
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
    SD_HIP(hipMalloc(&kst, sizeof(KrylovDev)));
    SD_HIP(hipMemsetAsync(kst, 0, sizeof(KrylovDev), c.stream));
    SD_HIP(hipHostMalloc(&kh, sizeof(KrylovHost), hipHostMallocMapped | hipHostMallocCoherent));
    SD_HIP(hipHostGetDevicePointer((void **)&khdev, kh, 0));
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
    if (kst) (void)hipFree(kst);
    if (kh) (void)hipHostFree(kh);
    kst = nullptr;
    kh = khdev = nullptr;
    for (void *p : {(void *)binv, (void *)bcol})
        if (p) (void)hipFree(p);
    binv = nullptr;
    bcol = nullptr;
    strong.release();
    strong_valid = false;
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
        if ((rc = ensure_ax(c))) return rc;
        hipLaunchKernelGGL(k_pair_partner, dim3((n + 3) / 4), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.Ax,
                           c.AxRow, c.partner, c.strength);
        c.partner_valid = true;
    }
    // without the column-major values the blocks read A_ji from the mirror
    // entry, which a symmetric pattern has
    if (!c.jac_symmetric && (rc = ensure_ax(c))) return rc;
    hipLaunchKernelGGL(k_pair_blocks, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag,
                       c.ax_valid ? c.Ax : nullptr, c.AxRow, c.partner, c.strength, 1e-3, binv, bcol);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// Whether the learned preconditioner serves the current matrix.  The
// network was trained on the Newton systems of running simulations, close to
// the identity; at a step size large for the configuration (the first
// attempts from h0 = 1e-6 at N = 16384) many diagonals move far from 1 and
// the pair blocks alone take fewer iterations (measured at N = 16384, m = 2:
// learned 40 without converging / 30 / 20 / 10 iterations against 37 / 24 /
// 18 / 9 for the pair blocks at h = 1e-6 / 5e-7 / 2.5e-7 / 1e-7, with 1034 /
// 490 / 248 / 90 strongly shifted diagonals; 6 against 8 at 5e-8 with 39, 3
// against 5 at 1e-9 with 2).  The count is the last full assembly's; a reuse
// of its values at another step size (stage II, h / 2) scales it by the
// step-size ratio, as every shift scales with h.
static bool learned_serves(const Context &c)
{
    if (c.learned_gate == -1) return true;
    const int gate = c.learned_gate >= 0 ? c.learned_gate : std::max(16, c.n / 256);
    double count = (double)c.diag_strong;
    if (c.diag_strong_h > 0.0 && c.jac_h != c.diag_strong_h) count *= c.jac_h / c.diag_strong_h;
    return count <= (double)gate;
}

int Fgmres::factor(Context &c)
{
    int rc = init(c);
    if (rc) return rc;
    strong_valid = false;
    c.precond_active = c.precond;
    if (c.precond == SDDDST_PRECOND_STRONG) return strong_factor(c);
    if (c.precond == SDDDST_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_jacobi_setup, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.diag, c.AxRow, dinv);
        SD_HIP(hipGetLastError());
        return SDDDST_OK;
    }
    if ((rc = pair_factor(c))) return rc;
    if (c.precond == SDDDST_PRECOND_LEARNED) {
        if (!learned_serves(c)) {
            c.precond_active = SDDDST_PRECOND_PAIR;
            return SDDDST_OK;
        }
        return spai_factor(c, *this);
    }
    return SDDDST_OK;
}

int Fgmres::apply(Context &c, const double *v, double *z)
{
    const int precond = c.precond_override >= 0 ? c.precond_override : c.precond_active;
    if (precond == SDDDST_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_jacobi_apply, dim3(nb256(n)), dim3(256), 0, c.stream, n, dinv, v, z);
    else if (precond == SDDDST_PRECOND_LEARNED)
        return spai_apply(c, *this, v, z);
    else
        hipLaunchKernelGGL(k_pair_apply, dim3(nb256(n)), dim3(256), 0, c.stream, n, binv, bcol, v, z);
    SD_HIP(hipGetLastError());
    if (precond == SDDDST_PRECOND_STRONG) return strong.apply(c, v, z);
    return SDDDST_OK;
}

// Threshold of the strong-row block: rows with |A_ii - 1| or an off-diagonal
// |A_ij| above it are solved exactly.  At 0.25 the first attempt of the
// N = 65536 point-defect run keeps about a quarter of the rows; FGMRES then
// converges in tens of iterations where the pair blocks alone stagnate.
constexpr double kStrongThreshold = 0.25;

int Fgmres::strong_factor(Context &c)
{
    if (strong_valid) return SDDDST_OK;
    int rc = pair_factor(c);
    if (rc) return rc;
    if ((rc = strong.factor(c, kStrongThreshold))) return rc;
    strong_valid = true;
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

// One restart cycle is driven by the device: the host enqueues a batch of
// iterations (z_k = M v_k, w = A z_k, two projection passes -- the first
// subtraction fused into the second, whose last workgroup also updates the
// Givens state -- and the next basis vector) and the cycle's closing kernels
// (back substitution with delta += Z y, the true residual g - A delta and
// its norm), then waits once.  Iteration kernels
// after the cycle stops, and the closing kernels while it still runs, return
// at once, so a solve that converges within a batch costs one wait.  Only an
// Arnoldi step whose second pass removed more than a quarter of ||w'|| comes
// back to the host, for an explicit norm.  After the first batch (as long
// as the context's recent solves) each batch is the number of iterations
// the cycle's mean reduction per iteration so far needs to bring the
// residual estimate to the target (at most kBatchMax), so a long solve
// (tens of iterations at a large step size) waits a few times instead of
// once every kBatch iterations.
constexpr int kBatch = 2, kBatchMax = 16;

static int next_batch(double res, double res_start, int k, double target)
{
    if (!(res > 0.0) || !(res_start > 0.0) || k <= 0 || res <= target) return 1;
    const double rate = std::log(res / res_start) / k;   // log of the mean reduction per iteration
    if (!(rate < -1e-3)) return kBatch;
    const double need = std::ceil(std::log(target / res) / rate);
    return (int)std::max(1.0, std::min((double)kBatchMax, need));
}

int Fgmres::solve(Context &c)
{
    const int N = n;
    const double tol = c.solver_tol;
    // a learned-preconditioned solve gets one restart cycle; the caller
    // retries with the pair blocks when it does not converge in it
    const int precond = c.precond_override >= 0 ? c.precond_override : c.precond_active;
    const int max_it = precond == SDDDST_PRECOND_LEARNED ? std::min(c.solver_max_iter, m) : c.solver_max_iter;
    int rc;
    // delta = 0, r = g: ||g||^2 into the device slot gs, v_0 = g / ||g||
    SD_HIP(hipMemsetAsync(c.delta, 0, sizeof(double) * N, c.stream));
    double *gs = hcol + 3 * (m + 2);
    double *rn2 = hcol + 3 * (m + 2) + 1;   // ||g - A delta||^2 of a cycle's end (device; H[0] on the host)
    if (c.gnorm2_valid) {
        // formed by the residual kernel that wrote g
        gs = c.gnorm2;
        c.gnorm2_valid = false;
    } else {
        launch_proj(c, *this, 0, V, const_cast<double *>(c.g), nullptr, gs, nullptr, 0, 1);
        SD_HIP(hipGetLastError());
    }
    c.last_iters = 0;
    int total_it = 0;
    double gnorm = -1.0, rel = 1.0, beta = 0.0;
    const double *h1 = hcol, *h2 = hcol + (m + 2);
    double *hsum = hcol + 2 * (m + 2);
    const int *stop = &kst->stop;
    for (int cycle = 0;; ++cycle) {
        const int kmax = std::min(m, max_it - total_it);
        if (kmax <= 0) break;
        // v_0 = r / ||r||: r = g in cycle 0, else the last cycle's true residual (in w)
        hipLaunchKernelGGL(k_cycle_start, dim3(nb256(N)), dim3(256), 0, c.stream, N, cycle == 0 ? c.g : w,
                           cycle == 0 ? gs : rn2, V, kst, khdev, cycle == 0 ? 1 : 0, kmax);
        SD_HIP(hipGetLastError());
        int k = 0;   // iterations of this cycle the host has seen completed
        bool ended = false;
        // the first batch as long as this context's recent solves (rounded),
        // later ones from the observed reduction (next_batch): iterations
        // past convergence return at once but still cost their launches
        int batch = std::max(1, std::min(m, (int)(iter_mean + 0.5)));
        const double res_start = cycle == 0 ? 1.0 : rel;
        while (!ended) {
            const int B = std::min(batch, kmax - k);
            batch = kBatch;
            for (int b = 0; b < B; ++b) {
                const int kk = k + b;
                double *vk = V + (size_t)kk * N, *zk = Z + (size_t)kk * N;
                if ((rc = apply(c, vk, zk))) return rc;
                if ((rc = launch_spmv(c, zk, w, stop))) return rc;
                launch_proj(c, *this, kk + 1, V, w, nullptr, const_cast<double *>(h1), hsum, 0, 0, nullptr, 0, stop);
                launch_proj(c, *this, kk + 1, V, w, h1, const_cast<double *>(h2), hsum, 1, 1, nullptr, 0, stop,
                            kGateIter, 0.5 * tol);
                hipLaunchKernelGGL(k_orth_scale_dev, dim3(nb256(N)), dim3(256), 0, c.stream, N, kk + 1, V, h2, w, kst,
                                   V + (size_t)(kk + 1) * N);
                SD_HIP(hipGetLastError());
            }
            // the cycle's end, when it came: delta += Z y and the true residual
            hipLaunchKernelGGL(k_combine_dev, dim3(nb256(N)), dim3(256), 0, c.stream, N, Z, kst, c.delta);
            bool fused = false;
            if ((rc = launch_spmv(c, c.delta, w, stop, kGateClose, c.g, this, rn2, Hdev, &fused))) return rc;
            if (!fused) launch_proj(c, *this, 0, V, w, nullptr, rn2, nullptr, 0, 1, Hdev, 0, stop, kGateClose);
            SD_HIP(hipGetLastError());
            if ((rc = host_wait(c))) return rc;
            const volatile KrylovHost *vkh = kh;
            KrylovHost st;
            st.gnorm = vkh->gnorm;
            st.res = vkh->res;
            st.k = vkh->k;
            st.stop = vkh->stop;
            if (gnorm < 0) gnorm = st.gnorm;
            SD_CHECK(st.k >= k && (st.k > k || st.stop != kRunning || B == 0), SDDDST_ERR_SOLVER,
                     "FGMRES: no progress in a batch (iteration %d, state %d)", k, st.stop);
            total_it += st.k - k;
            k = st.k;
            switch (st.stop) {
            case kZeroRhs:
                c.last_iters = 0;
                c.last_rel_res = 0.0;
                c.last_gnorm = 0.0;
                return SDDDST_OK;
            case kNeedNorm: {
                // iteration k: w -= V h2, its norm taken explicitly, then the
                // step resumed with it (the batch's later kernels returned)
                hipLaunchKernelGGL(k_orth, dim3(nb256(N)), dim3(256), 0, c.stream, N, k + 1, V, h2, w);
                SD_HIP(hipGetLastError());
                double hn;
                if ((rc = norm2(c, *this, w, &hn))) return rc;
                hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, c.stream, kst, khdev, hcol, m, k, 0.5 * tol,
                                   hn);
                if (hn > 0)
                    hipLaunchKernelGGL(k_scale_copy, dim3(nb256(N)), dim3(256), 0, c.stream, N, 1.0 / hn, w,
                                       V + (size_t)(k + 1) * N);
                SD_HIP(hipGetLastError());
                ++k;
                ++total_it;
                break;
            }
            case kRunning:
                batch = next_batch(st.res, res_start, k, 0.5 * tol);
                break;
            default:   // converged estimate, breakdown or budget: the closing kernels ran
                beta = sqrt(H[0]);
                rel = beta / gnorm;
                ended = true;
            }
        }
        if (rel <= tol || total_it >= max_it) break;
    }
    c.last_iters = total_it;
    c.last_rel_res = rel;
    c.last_gnorm = gnorm;
    iter_mean = 0.875 * iter_mean + 0.125 * total_it;
    if (!(c.last_rel_res <= tol)) {
        set_error("FGMRES: relative residual %.3e after %d iterations", c.last_rel_res, total_it);
        return SDDDST_ERR_SOLVER;
    }
    return SDDDST_OK;
}

}  // namespace sdamd
