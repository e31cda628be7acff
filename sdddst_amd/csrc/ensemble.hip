// Batched ensembles: many independent realizations of uneven size advanced
// in lock step on one device.  The reference runs one realization per
// process; each realization here follows Simulation::step
// (simulation.cpp:435-650) exactly as the single-system path does, with a
// constant (or zero) external stress, which takes the shortcut of
// simulation.cpp:75.
//
// One attempt = stage I (big step h), stage II (first half step), stage III
// (second half step), error control, commit; every kernel covers every
// realization, and per-realization flags on the device (active, accepted,
// capacity overflow) gate the work, so an attempt is a fixed launch sequence
// that is captured once as a HIP graph and replayed.  The linear solves are
// one persistent launch per solve: a workgroup per realization runs FGMRES
// with the pair-block preconditioner to the relative residual 1e-10 and,
// should it not converge, an LU factorisation with partial pivoting of the
// dense matrix in the same launch.
//
// HBM layout: realizations are concatenated (global dislocation index
// g = off[s] + local i); the CSC arrays of realization s start at nzoff[s]
// with the capacity cap[s] (regrown by the host when a count exceeds it),
// Ap of s at off[s] + s, the Krylov bases and dense workspaces at their own
// per-realization offsets.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "context.h"
#include "field.h"
#include "jac_core.h"
#include "pair_core.h"
#include "precond_core.h"
#include "step_core.h"

namespace sdamd {

namespace ens {

constexpr int kRestart = 30;       // FGMRES restart length
constexpr int kMaxIter = 300;      // FGMRES iterations before the LU fallback
constexpr int kDenseMax = 4096;    // largest realization with an LU workspace
enum { DISL = 0, BIG = 1, FS = 2, SS = 3, NSLOT = 4 };
enum { VINIT = 0, VINIT2 = 1, VSPEED = 2, VSPEED2 = 3, NVB = 4 };

struct SysDesc {
    int off, n, words, nb;          // nb: 64-wide tiles per side
    long long moff, nzoff, cap, voff, zoff, loff, poff;
    jac::CutParams cp;
};

struct SysState {
    double h, t, h_next, max_ratio;
    long long steps, failed, nnz, need, iters, solves, lu_fallbacks, solve_fail;
    int active, accept, ovf, selected;
};

}  // namespace ens

using ens::SysDesc;
using ens::SysState;

namespace {

__device__ __forceinline__ bool live(const SysState &s) { return s.active && !s.ovf; }

__global__ void k_e_copy(int nt, const int *__restrict__ sys, const SysState *__restrict__ st, int only_accepted,
                         const double *__restrict__ src, double *__restrict__ dst)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= nt) return;
    const SysState &q = st[sys[g]];
    if (!live(q) || (only_accepted && !q.accept)) return;
    dst[g] = src[g];
}

// Speeds of one realization's targets, one lane per target, sources staged
// through LDS in blocks of blockDim.x; each pair term is evaluated with the
// lower index first and accumulated in the reference's order
// (simulation.cpp:86-131), so the exact field gives the reference's bits.
template <int Mode>
__global__ void k_e_speeds(const int *__restrict__ blk, const SysDesc *__restrict__ D,
                           const SysState *__restrict__ st, int only_accepted, const double *__restrict__ x,
                           const double *__restrict__ y, const double *__restrict__ b, double sigma,
                           double *__restrict__ tol, double minPrecSqr, int update_tol, double *__restrict__ v)
{
    extern __shared__ double sm[];
    const int T = (int)blockDim.x;
    double *sx = sm, *sy = sm + T, *sb = sm + 2 * T;
    const int s = blk[2 * blockIdx.x], first = blk[2 * blockIdx.x + 1];
    const SysState &q = st[s];
    if (!live(q) || (only_accepted && !q.accept)) return;
    const int n = D[s].n, off = D[s].off;
    const int i = first + (int)threadIdx.x;
    const bool own = i < n;
    const double xi = own ? x[off + i] : 0.0, yi = own ? y[off + i] : 0.0, bI = own ? b[off + i] : 0.0;
    double acc = 0.0;
    double t = (own && update_tol) ? tol[off + i] : 0.0;
    for (int j0 = 0; j0 < n; j0 += T) {
        __syncthreads();
        const int jl = j0 + (int)threadIdx.x;
        if (jl < n) {
            sx[threadIdx.x] = x[off + jl];
            sy[threadIdx.x] = y[off + jl];
            sb[threadIdx.x] = b[off + jl];
        }
        __syncthreads();
        if (!own) continue;
        const int jend = min(T, n - j0);
        for (int jj = 0; jj < jend; ++jj) {
            const int j = j0 + jj;
            if (j == i) continue;
            double dx, dy, tmp;
            if (j < i) {
                dx = sdfield::normalize(sx[jj] - xi);
                dy = sdfield::normalize(sy[jj] - yi);
            } else {
                dx = sdfield::normalize(xi - sx[jj]);
                dy = sdfield::normalize(yi - sy[jj]);
            }
            double tau;
            if constexpr (Mode == SDDDST_FIELD_FAST) tau = sdfield::xy_fast(dx, dy);
            else tau = sdfield::xy(dx, dy);
            if (update_tol) stepf::tol_update(t, dx * dx + dy * dy, minPrecSqr);
            if (j < i) {
                tmp = sb[jj] * bI * tau;
                acc -= tmp;
            } else {
                tmp = bI * sb[jj] * tau;
                acc += tmp;
            }
        }
    }
    if (!own) return;
    acc += bI * sigma;
    v[off + i] = acc;
    if (update_tol) tol[off + i] = t;
}

__global__ void k_e_fill_minr2(int nt, const int *__restrict__ sys, const SysState *__restrict__ st,
                               int only_accepted, unsigned long long *__restrict__ minr2)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= nt) return;
    const SysState &q = st[sys[g]];
    if (!live(q) || (only_accepted && !q.accept)) return;
    minr2[g] = pairtile::kInfBits;
}

// The symmetric tiles of every realization (tile bodies shared with the
// single-system kernel); tl holds (realization, tile id) per workgroup.
template <int Mode>
__global__ __launch_bounds__(64) void k_e_pair_tiles(const int *__restrict__ tl, const SysDesc *__restrict__ D,
                                                     const SysState *__restrict__ st, int only_accepted,
                                                     const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ b, double *__restrict__ partial,
                                                     unsigned long long *__restrict__ minr2, double r2_keep)
{
    const int s = tl[2 * blockIdx.x];
    const SysState &q = st[s];
    if (!live(q) || (only_accepted && !q.accept)) return;
    int bi, bj;
    pairtile::tile_coords(tl[2 * blockIdx.x + 1], bi, bj);
    const SysDesc &d = D[s];
    __shared__ pairtile::TileLds L;
    pairtile::tile<Mode>(L, bi, bj, d.n, x + d.off, y + d.off, b + d.off, nullptr, partial + d.poff, minr2 + d.off,
                         r2_keep);
}

// Sum of a target's tile partials in tile order, tolerance update, stress.
__global__ void k_e_speeds_finish(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                  const SysState *__restrict__ st, int only_accepted,
                                  const double *__restrict__ partial, const unsigned long long *__restrict__ minr2,
                                  const double *__restrict__ b, double sigma, double *__restrict__ tol,
                                  double minPrecSqr, int update_tol, double *__restrict__ v)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= nt) return;
    const int s = sys[g];
    const SysState &q = st[s];
    if (!live(q) || (only_accepted && !q.accept)) return;
    const SysDesc &d = D[s];
    const int i = g - d.off;
    const double *ps = partial + d.poff;
    double acc = 0.0;
    int t = 0;
    for (; t + 8 <= d.nb; t += 8) {
        double p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = ps[(size_t)(t + u) * d.n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += p[u];
    }
    for (; t < d.nb; ++t) acc += ps[(size_t)t * d.n + i];
    if (update_tol) {
        double t = tol[g];
        const unsigned long long bits = minr2[g];
        if (bits != pairtile::kInfBits) stepf::tol_update(t, __longlong_as_double((long long)bits), minPrecSqr);
        tol[g] = t;
    }
    acc += b[g] * sigma;
    v[g] = acc;
}

// -- Jacobian (column bodies shared with the single-system path) -------------

struct ColView {
    int s, j;
};

__device__ __forceinline__ bool column_of(int g, int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                          const SysState *__restrict__ st, bool need_fit, ColView &cv)
{
    if (g >= nt) return false;
    cv.s = sys[g];
    const SysState &q = st[cv.s];
    if (!q.active) return false;
    if (need_fit && q.ovf) return false;
    cv.j = g - D[cv.s].off;
    return true;
}

__global__ __launch_bounds__(256) void k_e_jac_count(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                                     const SysState *__restrict__ st, const double *__restrict__ x,
                                                     const double *__restrict__ y, int *__restrict__ cnt,
                                                     unsigned long long *__restrict__ mask)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, false, cv)) return;
    const SysDesc &d = D[cv.s];
    jac::count_column(cv.j, d.n, d.words, x + d.off, y + d.off, d.cp, cnt + g, mask + d.moff + (size_t)cv.j * d.words);
}

// Exclusive scan of one realization's counts (workgroup per realization);
// a total above the capacity flags the realization for regrowth.
__global__ __launch_bounds__(1024) void k_e_scan(const SysDesc *__restrict__ D, SysState *__restrict__ st,
                                                 const int *__restrict__ cnt, int *__restrict__ Ap)
{
    __shared__ long long part[1024];
    const int s = (int)blockIdx.x;
    if (!st[s].active) return;
    const SysDesc &d = D[s];
    const int n = d.n;
    const int *c = cnt + d.off;
    int *ap = Ap + d.off + s;
    const int t = (int)threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int s0 = t * per, s1 = min(n, s0 + per);
    long long acc = 0;
    for (int k = s0; k < s1; ++k) acc += c[k];
    part[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const long long total = part[1023];
    const bool fits = total <= d.cap;
    long long run = part[t] - acc;
    if (fits)
        for (int k = s0; k < s1; ++k) {
            ap[k] = (int)run;
            run += c[k];
        }
    if (t == 0) {
        st[s].nnz = total;
        if (fits) {
            ap[n] = (int)total;
        } else {
            st[s].ovf = 1;
            st[s].need = total;
        }
    }
}

__global__ __launch_bounds__(256) void k_e_jac_fill(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                                    const SysState *__restrict__ st,
                                                    const unsigned long long *__restrict__ mask,
                                                    const int *__restrict__ Ap, int *__restrict__ Ai,
                                                    int *__restrict__ diag)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    jac::fill_column(cv.j, d.words, mask + d.moff + (size_t)cv.j * d.words, Ap[d.off + cv.s + cv.j], Ai + d.nzoff,
                     diag + g);
}

template <int Mode>
__global__ __launch_bounds__(256) void k_e_jac_values(int nt, const int *__restrict__ sys,
                                                      const SysDesc *__restrict__ D, const SysState *__restrict__ st,
                                                      const double *__restrict__ x, const double *__restrict__ y,
                                                      const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                      const int *__restrict__ diag, double *__restrict__ tau,
                                                      double *__restrict__ mult, int *__restrict__ mirror)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    jac::values_column<Mode>(cv.j, x + d.off, y + d.off, d.cp, Ap + d.off + cv.s, Ai + d.nzoff, diag + d.off,
                       tau + d.nzoff, mult + d.nzoff, mirror + d.nzoff, nullptr);
}

__global__ __launch_bounds__(256) void k_e_jac_raw(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                                   const SysState *__restrict__ st, double hscale,
                                                   const double *__restrict__ b, const int *__restrict__ Ap,
                                                   const int *__restrict__ Ai, const int *__restrict__ diag,
                                                   const double *__restrict__ tau, const double *__restrict__ mult,
                                                   const int *__restrict__ mirror, double *__restrict__ Ax,
                                                   int *__restrict__ flags)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    // exact zeros are kept in the pattern here: they do not change the solve
    jac::raw_column(cv.j, b + d.off, st[cv.s].h * hscale, Ap + d.off + cv.s, Ai + d.nzoff, diag + d.off,
                    tau + d.nzoff, mult + d.nzoff, mirror + d.nzoff, 0.0, Ax + d.nzoff, flags + cv.s);
}

__global__ __launch_bounds__(256) void k_e_jac_diag(int nt, const int *__restrict__ sys,
                                                    const SysDesc *__restrict__ D, const SysState *__restrict__ st,
                                                    const int *__restrict__ Ap, const int *__restrict__ diag,
                                                    double *__restrict__ Ax, double *__restrict__ dvec)
{
    ColView cv;
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    jac::diag_column(cv.j, Ap + d.off + cv.s, diag + d.off, Ax + d.nzoff, dvec + d.off);
}

__global__ __launch_bounds__(256) void k_e_jac_scale(int nt, const int *__restrict__ sys,
                                                     const SysDesc *__restrict__ D, const SysState *__restrict__ st,
                                                     const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                     const int *__restrict__ diag, const double *__restrict__ dvec,
                                                     double *__restrict__ Ax, double *__restrict__ AxRow)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    jac::scale_column(cv.j, Ap + d.off + cv.s, Ai + d.nzoff, diag + d.off, dvec + d.off, Ax + d.nzoff,
                      AxRow + d.nzoff);
}

__global__ __launch_bounds__(256) void k_e_partner(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                                                   const SysState *__restrict__ st, const int *__restrict__ Ap,
                                                   const int *__restrict__ Ai, const int *__restrict__ diag,
                                                   const double *__restrict__ Ax, const double *__restrict__ AxRow,
                                                   int *__restrict__ partner, double *__restrict__ strength)
{
    ColView cv;
    const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    pairblk::partner_row(cv.j, Ap + d.off + cv.s, Ai + d.nzoff, diag + d.off, Ax + d.nzoff, AxRow + d.nzoff,
                         partner + d.off, strength + d.off);
}

__global__ void k_e_blocks(int nt, const int *__restrict__ sys, const SysDesc *__restrict__ D,
                           const SysState *__restrict__ st, const int *__restrict__ Ap, const int *__restrict__ Ai,
                           const int *__restrict__ diag, const double *__restrict__ Ax,
                           const double *__restrict__ AxRow, const int *__restrict__ partner,
                           const double *__restrict__ strength, double *__restrict__ binv, int *__restrict__ bcol)
{
    ColView cv;
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (!column_of(g, nt, sys, D, st, true, cv)) return;
    const SysDesc &d = D[cv.s];
    pairblk::block_row(cv.j, Ap + d.off + cv.s, Ai + d.nzoff, diag + d.off, Ax + d.nzoff, AxRow + d.nzoff,
                       partner + d.off, strength + d.off, 1e-3, binv + 2 * (size_t)d.off, bcol + d.off);
}

__global__ void k_e_residual(int nt, const int *__restrict__ sys, const SysState *__restrict__ st, double hscale,
                             const double *__restrict__ xn, const double *__restrict__ xo,
                             const double *__restrict__ csp, const double *__restrict__ isp,
                             const double *__restrict__ dvec, double *__restrict__ g)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nt) return;
    const SysState &q = st[sys[i]];
    if (!live(q)) return;
    g[i] = stepf::residual(xn[i], xo[i], csp[i], isp[i], dvec[i], q.h * hscale);
}

__global__ void k_e_update(int nt, const int *__restrict__ sys, const SysState *__restrict__ st,
                           double *__restrict__ x, const double *__restrict__ delta)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nt) return;
    if (!live(st[sys[i]])) return;
    x[i] -= delta[i];
}

// -- persistent FGMRES, one workgroup per realization -------------------------

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *red, double *bcast)
{
    constexpr int NW = BS / 64;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < NW; ++k) s += red[k];
        *bcast = s;
    }
    __syncthreads();
    return *bcast;
}

// out = A v on the CSR view: groups of L lanes per row, 64 / L rows per
// wavefront in flight (L is a tuning parameter: short rows want small L).
template <int BS, int L>
__device__ __forceinline__ void block_spmv(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                           const double *__restrict__ Av, const double *__restrict__ v,
                                           double *__restrict__ out)
{
    constexpr int NW = BS / 64, G = 64 / L;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int sub = lane / L, sl = lane % L;
    for (int r0 = w * G; r0 < n; r0 += NW * G) {
        const int r = r0 + sub;
        double s = 0.0;
        if (r < n)
            for (int k = Ap[r] + sl; k < Ap[r + 1]; k += L) s += Av[k] * v[Ai[k]];
        for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (sl == 0 && r < n) out[r] = s;
    }
    __syncthreads();
}

// LU with partial pivoting of the dense matrix (column-major, n x n) and the
// solve of A x = rhs in place; false when a pivot is zero.
template <int BS>
__device__ bool block_dense_solve(int n, double *__restrict__ L, double *__restrict__ rhs, double *red,
                                  int *redi, double *bcast, int *bcasti)
{
    constexpr int NW = BS / 64;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = n;
        for (int i = k + (int)threadIdx.x; i < n; i += BS) {
            const double a = fabs(L[i + (size_t)k * n]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_down(best, o, 64);
            const int oi = __shfl_down(bi, o, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        __syncthreads();
        if (lane == 0) {
            red[w] = best;
            redi[w] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double b0 = red[0];
            int i0 = redi[0];
            for (int q = 1; q < NW; ++q)
                if (red[q] > b0 || (red[q] == b0 && redi[q] < i0)) {
                    b0 = red[q];
                    i0 = redi[q];
                }
            *bcast = b0;
            *bcasti = i0;
        }
        __syncthreads();
        const int p = *bcasti;
        if (!(*bcast > 0.0)) return false;
        if (p != k) {
            for (int c = (int)threadIdx.x; c < n; c += BS) {
                const double t = L[k + (size_t)c * n];
                L[k + (size_t)c * n] = L[p + (size_t)c * n];
                L[p + (size_t)c * n] = t;
            }
            if (threadIdx.x == 0) {
                const double t = rhs[k];
                rhs[k] = rhs[p];
                rhs[p] = t;
            }
        }
        __syncthreads();
        const double piv = L[k + (size_t)k * n];
        for (int i = k + 1 + (int)threadIdx.x; i < n; i += BS) L[i + (size_t)k * n] /= piv;
        __syncthreads();
        const int m = n - k - 1;
        for (long long t = threadIdx.x; t < (long long)m * m; t += BS) {
            const int i = k + 1 + (int)(t % m), c = k + 1 + (int)(t / m);
            L[i + (size_t)c * n] -= L[i + (size_t)k * n] * L[k + (size_t)c * n];
        }
        __syncthreads();
        for (int i = k + 1 + (int)threadIdx.x; i < n; i += BS) rhs[i] -= L[i + (size_t)k * n] * rhs[k];
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        if (threadIdx.x == 0) rhs[k] /= L[k + (size_t)k * n];
        __syncthreads();
        for (int i = (int)threadIdx.x; i < k; i += BS) rhs[i] -= L[i + (size_t)k * n] * rhs[k];
        __syncthreads();
    }
    return true;
}

template <int BS, int L>
__global__ __launch_bounds__(BS) void k_e_fgmres(const SysDesc *__restrict__ D, SysState *__restrict__ st,
                                                 const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                 const double *__restrict__ Ax, const double *__restrict__ AxRow,
                                                 const double *__restrict__ binv, const int *__restrict__ bcol,
                                                 const double *__restrict__ gall, double *__restrict__ dall,
                                                 double *__restrict__ Vall, double *__restrict__ Zall,
                                                 double *__restrict__ dense, double tol)
{
    constexpr int NW = BS / 64;
    constexpr int M = ens::kRestart;
    __shared__ double H[(M + 1) * M];
    __shared__ double cs[M], sn[M], sv[M + 1], yv[M], hp[M + 1];
    __shared__ double red[NW * (M + 1)];
    __shared__ int redi[NW];
    __shared__ double bcast;
    __shared__ int bcasti, stop;
    const int s = (int)blockIdx.x;
    if (!live(st[s])) return;
    const SysDesc d = D[s];
    const int n = d.n;
    const int *ap = Ap + d.off + s;
    const int *ai = Ai + d.nzoff;
    const double *ar = AxRow + d.nzoff;
    const double *bi = binv + 2 * (size_t)d.off;
    const int *bc = bcol + d.off;
    const double *g = gall + d.off;
    double *x = dall + d.off;
    double *V = Vall + d.voff;
    double *Z = Zall + d.zoff;
    const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;

    for (int i = tid; i < n; i += BS) x[i] = 0.0;
    double acc = 0;
    for (int i = tid; i < n; i += BS) acc += g[i] * g[i];
    const double gnorm = sqrt(block_sum<BS>(acc, red, &bcast));
    int total = 0;
    if (gnorm == 0.0) {
        if (tid == 0) st[s].solves += 1;
        return;
    }
    double rel = 1.0;
    for (int cycle = 0; total < ens::kMaxIter; ++cycle) {
        // V0 = g - A x
        if (cycle == 0) {
            for (int i = tid; i < n; i += BS) V[i] = g[i];
            __syncthreads();
        } else {
            block_spmv<BS, L>(n, ap, ai, ar, x, V);
            for (int i = tid; i < n; i += BS) V[i] = g[i] - V[i];
            __syncthreads();
        }
        acc = 0;
        for (int i = tid; i < n; i += BS) acc += V[i] * V[i];
        const double beta = sqrt(block_sum<BS>(acc, red, &bcast));
        rel = beta / gnorm;
        if (rel <= tol) break;
        for (int i = tid; i < n; i += BS) V[i] /= beta;
        if (tid <= M) sv[tid] = tid == 0 ? beta : 0.0;
        if (tid == 0) stop = 0;
        __syncthreads();
        int k = 0;
        for (; k < M && total < ens::kMaxIter; ++k, ++total) {
            const double *vk = V + (size_t)k * n;
            double *zk = Z + (size_t)k * n;
            double *wv = V + (size_t)(k + 1) * n;
            for (int i = tid; i < n; i += BS) {
                const int j = bc[i];
                double r = bi[2 * i] * vk[i];
                if (j >= 0) r += bi[2 * i + 1] * vk[j];
                zk[i] = r;
            }
            __syncthreads();
            block_spmv<BS, L>(n, ap, ai, ar, zk, wv);
            // classical Gram-Schmidt, twice
            if (tid <= k) H[tid * M + k] = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                for (int v = 0; v <= k; ++v) {
                    double a = 0;
                    const double *vv = V + (size_t)v * n;
                    for (int i = tid; i < n; i += BS) a += vv[i] * wv[i];
                    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
                    if (lane == 0) red[w * (M + 1) + v] = a;
                }
                __syncthreads();
                if (tid <= k) {
                    double a = 0;
                    for (int q = 0; q < NW; ++q) a += red[q * (M + 1) + tid];
                    hp[tid] = a;
                    H[tid * M + k] += a;
                }
                __syncthreads();
                for (int i = tid; i < n; i += BS) {
                    double a = wv[i];
                    for (int v = 0; v <= k; ++v) a -= hp[v] * V[(size_t)v * n + i];
                    wv[i] = a;
                }
                __syncthreads();
            }
            acc = 0;
            for (int i = tid; i < n; i += BS) acc += wv[i] * wv[i];
            const double hn = sqrt(block_sum<BS>(acc, red, &bcast));
            if (hn > 0)
                for (int i = tid; i < n; i += BS) wv[i] /= hn;
            if (tid == 0) {
                double hh[M + 1];
                for (int i = 0; i <= k; ++i) hh[i] = H[i * M + k];
                hh[k + 1] = hn;
                for (int i = 0; i < k; ++i) {
                    const double t = cs[i] * hh[i] + sn[i] * hh[i + 1];
                    hh[i + 1] = -sn[i] * hh[i] + cs[i] * hh[i + 1];
                    hh[i] = t;
                }
                const double den = hypot(hh[k], hh[k + 1]);
                cs[k] = den > 0 ? hh[k] / den : 1.0;
                sn[k] = den > 0 ? hh[k + 1] / den : 0.0;
                hh[k] = den;
                sv[k + 1] = -sn[k] * sv[k];
                sv[k] = cs[k] * sv[k];
                for (int i = 0; i <= k; ++i) H[i * M + k] = hh[i];
                stop = (fabs(sv[k + 1]) / gnorm <= 0.5 * tol || hn == 0.0) ? 1 : 0;
            }
            __syncthreads();
            if (stop) {
                ++k;
                ++total;
                break;
            }
        }
        if (tid == 0) {
            for (int i = k - 1; i >= 0; --i) {
                double a = sv[i];
                for (int j = i + 1; j < k; ++j) a -= H[i * M + j] * yv[j];
                yv[i] = a / H[i * M + i];
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += BS) {
            double a = x[i];
            for (int j = 0; j < k; ++j) a += yv[j] * Z[(size_t)j * n + i];
            x[i] = a;
        }
        __syncthreads();
    }
    // true residual
    block_spmv<BS, L>(n, ap, ai, ar, x, V);
    acc = 0;
    for (int i = tid; i < n; i += BS) {
        const double r = g[i] - V[i];
        acc += r * r;
    }
    rel = sqrt(block_sum<BS>(acc, red, &bcast)) / gnorm;
    if (tid == 0) {
        st[s].iters += total;
        st[s].solves += 1;
    }
    if (rel <= tol) return;
    // fallback: dense LU of A (CSC values) in this workgroup
    if (d.loff < 0) {
        if (tid == 0) st[s].solve_fail += 1;
        return;
    }
    double *LU = dense + d.loff;
    for (long long t = tid; t < (long long)n * n; t += BS) LU[t] = 0.0;
    __syncthreads();
    for (int c = 0; c < n; ++c)
        for (int k = ap[c] + tid; k < ap[c + 1]; k += BS) LU[ai[k] + (size_t)c * n] = Ax[d.nzoff + k];
    for (int i = tid; i < n; i += BS) x[i] = g[i];
    __syncthreads();
    const bool ok = block_dense_solve<BS>(n, LU, x, red, redi, &bcast, &bcasti);
    if (tid == 0) {
        st[s].lu_fallbacks += 1;
        if (!ok) st[s].solve_fail += 1;
    }
}

// Error control per realization (simulation.cpp:398-405, 516, 643-649):
// max ratio and its index, acceptance, next step size, tolerance reset.
__global__ __launch_bounds__(1024) void k_e_error(const SysDesc *__restrict__ D, SysState *__restrict__ st,
                                                  const double *__restrict__ xbig, const double *__restrict__ xss,
                                                  double *__restrict__ tol, double minPrecSqr)
{
    __shared__ double smax[1024];
    __shared__ int sidx[1024];
    const int s = (int)blockIdx.x;
    if (!live(st[s])) return;
    const SysDesc &d = D[s];
    const int t = (int)threadIdx.x;
    double best = 0.0;
    int bidx = 0x7fffffff;
    for (int i = t; i < d.n; i += 1024) {
        const int g = d.off + i;
        const double e = __builtin_fabs(xbig[g] - xss[g]);
        if (0.0 < e) {
            const double r = e * e / tol[g];
            if (r > best) {
                best = r;
                bidx = i;
            }
        }
        tol[g] = minPrecSqr;
    }
    smax[t] = best;
    sidx[t] = bidx;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            const double ov = smax[t + o];
            const int oi = sidx[t + o];
            if (ov > smax[t] || (ov == smax[t] && oi < sidx[t])) {
                smax[t] = ov;
                sidx[t] = oi;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const double m = smax[0];
        st[s].max_ratio = m;
        st[s].selected = (m > 0.0) ? sidx[0] : 0;
        st[s].accept = m < 1.0 ? 1 : 0;
        st[s].h_next = stepf::new_step_size(m, st[s].h, -1.0);
    }
}

// Commit of the attempt (simulation.cpp:518-641): time, counters, step size;
// a realization that overflowed its Jacobian capacity keeps its state and
// repeats the attempt after the host regrows the arrays.
__global__ void k_e_commit(int ns, SysState *__restrict__ st, long long target)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= ns) return;
    SysState &q = st[s];
    if (!q.active) return;
    if (q.ovf) return;
    if (q.accept) {
        q.t += q.h;
        q.steps += 1;
    } else {
        q.failed += 1;
    }
    q.h = q.h_next;
    q.accept = 0;
    if (q.steps >= target) q.active = 0;
}

__global__ void k_e_activate(int ns, SysState *__restrict__ st, long long target)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= ns) return;
    st[s].active = st[s].steps < target ? 1 : 0;
    st[s].ovf = 0;
}


}  // namespace

// ---------------------------------------------------------------------------

struct Ensemble {
    int S = 0, NT = 0;
    sdddst_params p{};
    double sigma = 0;
    long long target = 0;
    hipStream_t stream = nullptr;
    std::vector<SysDesc> hd;
    SysDesc *D = nullptr;
    SysState *st = nullptr, *hst = nullptr;
    int *sys = nullptr;
    double *x[ens::NSLOT] = {}, *y = nullptr, *b = nullptr, *v[ens::NVB] = {}, *tol = nullptr;
    double *g = nullptr, *delta = nullptr, *dvec = nullptr;
    int *cnt = nullptr, *Ap = nullptr, *diag = nullptr, *flags = nullptr;
    unsigned long long *mask = nullptr;
    int *Ai = nullptr, *mirror = nullptr;
    double *Ax = nullptr, *AxRow = nullptr, *tau = nullptr, *mult = nullptr;
    int *partner = nullptr, *bcol = nullptr;
    double *strength = nullptr, *binv = nullptr;
    double *V = nullptr, *Z = nullptr, *dense = nullptr;
    long long nzTotal = 0, vTotal = 0, zTotal = 0, lTotal = 0, mTotal = 0;
    double *partial = nullptr;
    unsigned long long *minr2 = nullptr;
    long long pTotal = 0;
    // launch configuration (autotuned): speedsBlock 0 = symmetric tiles,
    // else one lane per target in workgroups of that size
    int speedsBlock = 0, solverBlock = 256, spmvLanes = 8;
    int *blk = nullptr, *tl = nullptr;
    int nblk = 0, blkFor = 0, ntiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long attempts = 0, regrows = 0;
};

namespace {

template <class T>
int ealloc(T **p, long long count)
{
    if (*p) SD_HIP(hipFree(*p));
    *p = nullptr;
    if (count < 1) count = 1;
    SD_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)count));
    return SDDDST_OK;
}

inline unsigned nb(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// expected entries of one column: pi (7.07 rc)^2 N = 157 m^2, with margin
long long default_cap(int n, double m)
{
    const double eff = m >= 1. / 12. * sqrt(2.0 * n) ? 1e20 : m;
    const double per = 400.0 * eff * eff + 64.0;
    const double c = std::min((double)n, per) * (double)n;
    return (long long)c;
}

int ens_layout(Ensemble &e)
{
    long long nz = 0, vo = 0, zo = 0, lo = 0, mo = 0, po = 0;
    for (int s = 0; s < e.S; ++s) {
        SysDesc &d = e.hd[s];
        d.poff = po;
        po += (long long)d.nb * d.n;
        d.moff = mo;
        mo += (long long)d.n * d.words;
        d.nzoff = nz;
        nz += d.cap;
        d.voff = vo;
        vo += (long long)(ens::kRestart + 1) * d.n;
        d.zoff = zo;
        zo += (long long)ens::kRestart * d.n;
        if (d.n <= ens::kDenseMax) {
            d.loff = lo;
            lo += (long long)d.n * d.n;
        } else {
            d.loff = -1;
        }
    }
    int rc;
    const bool grow_nz = nz != e.nzTotal;
    if (grow_nz) {
        if ((rc = ealloc(&e.Ai, nz)) || (rc = ealloc(&e.mirror, nz)) || (rc = ealloc(&e.Ax, nz)) ||
            (rc = ealloc(&e.AxRow, nz)) || (rc = ealloc(&e.tau, nz)) || (rc = ealloc(&e.mult, nz)))
            return rc;
        e.nzTotal = nz;
    }
    if (mo != e.mTotal) {
        if ((rc = ealloc(&e.mask, mo))) return rc;
        e.mTotal = mo;
    }
    if (vo != e.vTotal) {
        if ((rc = ealloc(&e.V, vo)) || (rc = ealloc(&e.Z, zo))) return rc;
        e.vTotal = vo;
        e.zTotal = zo;
    }
    if (lo != e.lTotal) {
        if ((rc = ealloc(&e.dense, lo))) return rc;
        e.lTotal = lo;
    }
    if (po != e.pTotal) {
        if ((rc = ealloc(&e.partial, po))) return rc;
        e.pTotal = po;
    }
    SD_HIP(hipMemcpyAsync(e.D, e.hd.data(), sizeof(SysDesc) * e.S, hipMemcpyHostToDevice, e.stream));
    return SDDDST_OK;
}

int build_tiles(Ensemble &e)
{
    std::vector<int> h;
    for (int s = 0; s < e.S; ++s) {
        const long long nt = (long long)e.hd[s].nb * (e.hd[s].nb + 1) / 2;
        for (long long t = 0; t < nt; ++t) {
            h.push_back(s);
            h.push_back((int)t);
        }
    }
    SD_CHECK(h.size() / 2 < (1u << 31), SDDDST_ERR_ARG, "too many tiles");
    int rc = ealloc(&e.tl, (long long)h.size());
    if (rc) return rc;
    SD_HIP(hipMemcpyAsync(e.tl, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, e.stream));
    SD_HIP(hipStreamSynchronize(e.stream));
    e.ntiles = (int)(h.size() / 2);
    return SDDDST_OK;
}

int build_blocks(Ensemble &e, int T)
{
    std::vector<int> h;
    for (int s = 0; s < e.S; ++s)
        for (int f = 0; f < e.hd[s].n; f += T) {
            h.push_back(s);
            h.push_back(f);
        }
    int rc = ealloc(&e.blk, (long long)h.size());
    if (rc) return rc;
    SD_HIP(hipMemcpyAsync(e.blk, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, e.stream));
    SD_HIP(hipStreamSynchronize(e.stream));
    e.nblk = (int)(h.size() / 2);
    e.blkFor = T;
    return SDDDST_OK;
}

int launch_speeds(Ensemble &e, int slot, int buf, int update_tol, int only_accepted)
{
    const double mp2 = e.p.precision * e.p.precision;
    if (e.speedsBlock == 0) {
        hipLaunchKernelGGL(k_e_fill_minr2, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, only_accepted,
                           e.minr2);
        if (e.p.field_mode == SDDDST_FIELD_FAST)
            hipLaunchKernelGGL(k_e_pair_tiles<SDDDST_FIELD_FAST>, dim3(e.ntiles), dim3(64), 0, e.stream, e.tl, e.D,
                               e.st, only_accepted, e.x[slot], e.y, e.b, e.partial, e.minr2,
                               pairtile::tolerance_r2(mp2));
        else
            hipLaunchKernelGGL(k_e_pair_tiles<SDDDST_FIELD_EXACT>, dim3(e.ntiles), dim3(64), 0, e.stream, e.tl, e.D,
                               e.st, only_accepted, e.x[slot], e.y, e.b, e.partial, e.minr2,
                               pairtile::tolerance_r2(mp2));
        hipLaunchKernelGGL(k_e_speeds_finish, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st,
                           only_accepted, e.partial, e.minr2, e.b, e.sigma, e.tol, mp2, update_tol, e.v[buf]);
        SD_HIP(hipGetLastError());
        return SDDDST_OK;
    }
    const size_t shm = sizeof(double) * 3 * e.speedsBlock;
    if (e.p.field_mode == SDDDST_FIELD_FAST)
        hipLaunchKernelGGL(k_e_speeds<SDDDST_FIELD_FAST>, dim3(e.nblk), dim3(e.speedsBlock), shm, e.stream, e.blk, e.D,
                           e.st, only_accepted, e.x[slot], e.y, e.b, e.sigma, e.tol, mp2, update_tol, e.v[buf]);
    else
        hipLaunchKernelGGL(k_e_speeds<SDDDST_FIELD_EXACT>, dim3(e.nblk), dim3(e.speedsBlock), shm, e.stream, e.blk,
                           e.D, e.st, only_accepted, e.x[slot], e.y, e.b, e.sigma, e.tol, mp2, update_tol, e.v[buf]);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// Jacobian of the positions in slot with step hscale*h; build: pattern too.
int launch_jacobian(Ensemble &e, int slot, double hscale, bool build)
{
    const unsigned wb = nb(e.NT, 4);
    if (build) {
        hipLaunchKernelGGL(k_e_jac_count, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.x[slot], e.y,
                           e.cnt, e.mask);
        hipLaunchKernelGGL(k_e_scan, dim3(e.S), dim3(1024), 0, e.stream, e.D, e.st, e.cnt, e.Ap);
        hipLaunchKernelGGL(k_e_jac_fill, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.mask, e.Ap, e.Ai,
                           e.diag);
        if (e.p.field_mode == SDDDST_FIELD_FAST)
            hipLaunchKernelGGL(k_e_jac_values<SDDDST_FIELD_FAST>, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D,
                               e.st, e.x[slot], e.y, e.Ap, e.Ai, e.diag, e.tau, e.mult, e.mirror);
        else
            hipLaunchKernelGGL(k_e_jac_values<SDDDST_FIELD_EXACT>, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D,
                               e.st, e.x[slot], e.y, e.Ap, e.Ai, e.diag, e.tau, e.mult, e.mirror);
    }
    hipLaunchKernelGGL(k_e_jac_raw, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, hscale, e.b, e.Ap, e.Ai,
                       e.diag, e.tau, e.mult, e.mirror, e.Ax, e.flags);
    hipLaunchKernelGGL(k_e_jac_diag, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.Ap, e.diag,
                       e.Ax, e.dvec);
    hipLaunchKernelGGL(k_e_jac_scale, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.Ap, e.Ai, e.diag,
                       e.dvec, e.Ax, e.AxRow);
    hipLaunchKernelGGL(k_e_partner, dim3(wb), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.Ap, e.Ai, e.diag,
                       e.Ax, e.AxRow, e.partner, e.strength);
    hipLaunchKernelGGL(k_e_blocks, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.D, e.st, e.Ap, e.Ai,
                       e.diag, e.Ax, e.AxRow, e.partner, e.strength, e.binv, e.bcol);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

template <int BS, int L>
void solve_variant(Ensemble &e, double tol)
{
    hipLaunchKernelGGL((k_e_fgmres<BS, L>), dim3(e.S), dim3(BS), 0, e.stream, e.D, e.st, e.Ap, e.Ai, e.Ax, e.AxRow,
                       e.binv, e.bcol, e.g, e.delta, e.V, e.Z, e.dense, tol);
}

template <int BS>
void solve_lanes(Ensemble &e, double tol)
{
    switch (e.spmvLanes) {
    case 4: solve_variant<BS, 4>(e, tol); break;
    case 16: solve_variant<BS, 16>(e, tol); break;
    case 64: solve_variant<BS, 64>(e, tol); break;
    default: solve_variant<BS, 8>(e, tol); break;
    }
}

int launch_solve(Ensemble &e)
{
    const double tol = 1e-10;
    switch (e.solverBlock) {
    case 512: solve_lanes<512>(e, tol); break;
    case 1024: solve_lanes<1024>(e, tol); break;
    default: solve_lanes<256>(e, tol); break;
    }
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// integrateSlots (simulation.cpp:62-84) for every live realization:
// Jacobian at newSlot, ic Newton iterations; iteration 0 takes the
// constant-stress shortcut (v_new := v_old).
int integrate(Ensemble &e, double hscale, int newSlot, int oldSlot, bool useSpeed2, bool calcInitSpeed, bool reuse)
{
    int rc;
    if ((rc = launch_jacobian(e, newSlot, hscale, !reuse))) return rc;
    const int isp = useSpeed2 ? ens::VINIT2 : ens::VINIT;
    const int cspFull = useSpeed2 ? ens::VSPEED2 : ens::VSPEED;
    for (int it = 0; it < e.p.iteration_count; ++it) {
        int csp = cspFull;
        if (it == 0) {
            if (calcInitSpeed && (rc = launch_speeds(e, oldSlot, isp, 1, 0))) return rc;
            csp = isp;
        } else {
            if ((rc = launch_speeds(e, newSlot, csp, 1, 0))) return rc;
        }
        hipLaunchKernelGGL(k_e_residual, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, hscale,
                           e.x[newSlot], e.x[oldSlot], e.v[csp], e.v[isp], e.dvec, e.g);
        SD_HIP(hipGetLastError());
        if ((rc = launch_solve(e))) return rc;
        hipLaunchKernelGGL(k_e_update, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, e.x[newSlot],
                           e.delta);
        SD_HIP(hipGetLastError());
    }
    return SDDDST_OK;
}

// One attempt of every active realization (Simulation::step).
int enqueue_attempt(Ensemble &e)
{
    int rc;
    const double mp2 = e.p.precision * e.p.precision;
    // stage I: bigStep from dislocations
    hipLaunchKernelGGL(k_e_copy, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 0, e.x[ens::DISL],
                       e.x[ens::BIG]);
    if ((rc = integrate(e, 1.0, ens::BIG, ens::DISL, false, false, false))) return rc;
    // stage II: first small step, same positions as stage I
    hipLaunchKernelGGL(k_e_copy, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 0, e.x[ens::DISL],
                       e.x[ens::FS]);
    if ((rc = integrate(e, 0.5, ens::FS, ens::DISL, false, false, true))) return rc;
    // stage III: second small step from the first
    hipLaunchKernelGGL(k_e_copy, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 0, e.x[ens::FS],
                       e.x[ens::SS]);
    if ((rc = integrate(e, 0.5, ens::SS, ens::FS, true, true, false))) return rc;
    // error control, acceptance, post-step speeds, commit
    hipLaunchKernelGGL(k_e_error, dim3(e.S), dim3(1024), 0, e.stream, e.D, e.st, e.x[ens::BIG], e.x[ens::SS], e.tol,
                       mp2);
    hipLaunchKernelGGL(k_e_copy, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 1, e.x[ens::SS],
                       e.x[ens::DISL]);
    if ((rc = launch_speeds(e, ens::DISL, ens::VINIT, 0, 1))) return rc;
    hipLaunchKernelGGL(k_e_commit, dim3(nb(e.S)), dim3(256), 0, e.stream, e.S, e.st, e.target);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

void drop_graph(Ensemble &e)
{
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
}

int capture(Ensemble &e)
{
    drop_graph(e);
    SD_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_attempt(e);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(e.stream, &gr);
    if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    SD_HIP(ec);
    e.graph = gr;
    SD_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    return SDDDST_OK;
}

int read_state(Ensemble &e)
{
    SD_HIP(hipMemcpyAsync(e.hst, e.st, sizeof(SysState) * e.S, hipMemcpyDeviceToHost, e.stream));
    SD_HIP(hipStreamSynchronize(e.stream));
    return SDDDST_OK;
}

// Regrow the CSC capacity of realizations whose count overflowed.
int regrow(Ensemble &e, bool &changed)
{
    changed = false;
    for (int s = 0; s < e.S; ++s) {
        if (!e.hst[s].ovf) continue;
        const long long need = e.hst[s].need;
        e.hd[s].cap = std::max(need + need / 4 + 64, e.hd[s].cap * 2);
        e.hd[s].cap = std::min(e.hd[s].cap, (long long)e.hd[s].n * e.hd[s].n);
        changed = true;
    }
    if (!changed) return SDDDST_OK;
    e.regrows++;
    int rc = ens_layout(e);
    if (rc) return rc;
    return SDDDST_OK;
}

}  // namespace

void ensemble_free(Ensemble *e)
{
    if (!e) return;
    drop_graph(*e);
    void *ptrs[] = {e->D,    e->st,   e->sys,     e->y,      e->b,     e->tol,  e->g,       e->delta,
                    e->dvec, e->cnt,  e->Ap,      e->diag,   e->flags, e->mask, e->Ai,      e->mirror,
                    e->Ax,   e->AxRow, e->tau,    e->mult,   e->partner, e->bcol, e->strength,
                    e->binv, e->V,    e->Z,       e->dense,  e->blk,  e->tl,   e->partial, e->minr2};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (double *p : e->x)
        if (p) (void)hipFree(p);
    for (double *p : e->v)
        if (p) (void)hipFree(p);
    if (e->hst) (void)hipHostFree(e->hst);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int ensemble_create(int S, const int *n, const double *x, const double *y, const double *b, const sdddst_params *p,
                    double sigma, double h0, Ensemble **out)
{
    SD_CHECK(S >= 1 && n && x && y && b && p && out, SDDDST_ERR_ARG, "bad ensemble arguments");
    SD_CHECK(p->precision > 0 && p->iteration_count >= 1, SDDDST_ERR_ARG, "bad parameters");
    SD_CHECK(p->field_mode == SDDDST_FIELD_FAST || p->field_mode == SDDDST_FIELD_EXACT, SDDDST_ERR_ARG,
             "ensembles support the fast and exact field modes");
    SD_CHECK(h0 > 0, SDDDST_ERR_ARG, "initial step must be positive");
    Ensemble *e = new Ensemble();
    e->S = S;
    e->p = *p;
    e->sigma = sigma;
    e->hd.resize(S);
    long long nt = 0;
    for (int s = 0; s < S; ++s) {
        if (n[s] < 2 || n[s] > (1 << 20)) {
            delete e;
            set_error("realization %d has %d dislocations (2 .. 2^20 supported)", s, n[s]);
            return SDDDST_ERR_ARG;
        }
        SysDesc &d = e->hd[s];
        d.off = (int)nt;
        d.n = n[s];
        d.words = (n[s] + 63) / 64;
        d.nb = d.words;
        double m = p->cutoff_multiplier;
        if (m >= 1. / 12. * sqrt(2.0 * double(n[s]))) m = 1e20;
        const double rc = 1. / sqrt((double)n[s]) * m;
        d.cp = jac::make_cut_params(rc, rc * rc, 1. / (rc * rc));
        d.cap = default_cap(n[s], p->cutoff_multiplier);
        nt += n[s];
    }
    if (nt > (1LL << 30)) {
        delete e;
        set_error("ensemble too large");
        return SDDDST_ERR_ARG;
    }
    e->NT = (int)nt;
    const int NT = e->NT;
    int rc;
    auto fail = [&](int code) {
        ensemble_free(e);
        return code;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        set_error("hipStreamCreate failed");
        return SDDDST_ERR_HIP;
    }
    if ((rc = ealloc(&e->D, S)) || (rc = ealloc(&e->st, S)) || (rc = ealloc(&e->sys, NT)) || (rc = ealloc(&e->y, NT)) ||
        (rc = ealloc(&e->b, NT)) || (rc = ealloc(&e->tol, NT)) || (rc = ealloc(&e->g, NT)) ||
        (rc = ealloc(&e->delta, NT)) || (rc = ealloc(&e->dvec, NT)) || (rc = ealloc(&e->cnt, NT)) ||
        (rc = ealloc(&e->Ap, (long long)NT + S)) || (rc = ealloc(&e->diag, NT)) || (rc = ealloc(&e->flags, S)) ||
        (rc = ealloc(&e->partner, NT)) || (rc = ealloc(&e->bcol, NT)) || (rc = ealloc(&e->strength, NT)) ||
        (rc = ealloc(&e->binv, 2LL * NT)) || (rc = ealloc(&e->minr2, NT)))
        return fail(rc);
    for (int k = 0; k < ens::NSLOT; ++k)
        if ((rc = ealloc(&e->x[k], NT))) return fail(rc);
    for (int k = 0; k < ens::NVB; ++k)
        if ((rc = ealloc(&e->v[k], NT))) return fail(rc);
    if (hipHostMalloc((void **)&e->hst, sizeof(SysState) * S, hipHostMallocDefault) != hipSuccess)
        return fail(SDDDST_ERR_NOMEM);
    if ((rc = ens_layout(*e))) return fail(rc);
    std::vector<int> hsys(NT);
    for (int s = 0; s < S; ++s)
        for (int i = 0; i < n[s]; ++i) hsys[e->hd[s].off + i] = s;
    std::vector<SysState> h(S);
    for (int s = 0; s < S; ++s) {
        memset(&h[s], 0, sizeof(SysState));
        h[s].h = h0;
        h[s].active = 1;
    }
    std::vector<double> tol0(NT, p->precision * p->precision);
    auto up = [&](void *dst, const void *src, size_t bytes) -> int {
        SD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
        return SDDDST_OK;
    };
    if ((rc = up(e->sys, hsys.data(), sizeof(int) * NT)) || (rc = up(e->x[ens::DISL], x, sizeof(double) * NT)) ||
        (rc = up(e->y, y, sizeof(double) * NT)) || (rc = up(e->b, b, sizeof(double) * NT)) ||
        (rc = up(e->tol, tol0.data(), sizeof(double) * NT)) || (rc = up(e->st, h.data(), sizeof(SysState) * S)))
        return fail(rc);
    if ((rc = build_tiles(*e))) return fail(rc);
    // the first step's speeds (simulation.cpp:450-456): tolerance updated
    if ((rc = launch_speeds(*e, ens::DISL, ens::VINIT, 1, 0))) return fail(rc);
    SD_HIP(hipStreamSynchronize(e->stream));
    *out = e;
    return SDDDST_OK;
}

int ensemble_set_tiles(Ensemble &e, int speeds_block, int solver_block, int spmv_lanes)
{
    SD_CHECK(speeds_block == 0 || speeds_block == 64 || speeds_block == 128 || speeds_block == 256, SDDDST_ERR_ARG,
             "speeds block must be 0 (symmetric tiles), 64, 128 or 256");
    SD_CHECK(solver_block == 256 || solver_block == 512 || solver_block == 1024, SDDDST_ERR_ARG,
             "solver block must be 256, 512 or 1024");
    SD_CHECK(spmv_lanes == 4 || spmv_lanes == 8 || spmv_lanes == 16 || spmv_lanes == 64, SDDDST_ERR_ARG,
             "lanes per row must be 4, 8, 16 or 64");
    if (speeds_block > 0 && speeds_block != e.blkFor) {
        int rc = build_blocks(e, speeds_block);
        if (rc) return rc;
    }
    e.speedsBlock = speeds_block;
    e.solverBlock = solver_block;
    e.spmvLanes = spmv_lanes;
    drop_graph(e);
    return SDDDST_OK;
}

// Attempts until every realization has `target` successful steps or
// max_attempts ran; graph-captured when use_graph.
int ensemble_run(Ensemble &e, long long target, long long max_attempts, bool use_graph, long long *done_attempts)
{
    int rc;
    if (target != e.target) drop_graph(e);
    e.target = target;
    hipLaunchKernelGGL(k_e_activate, dim3(nb(e.S)), dim3(256), 0, e.stream, e.S, e.st, target);
    SD_HIP(hipGetLastError());
    long long done = 0;
    while (done < max_attempts) {
        if ((rc = read_state(e))) return rc;
        bool any = false;
        for (int s = 0; s < e.S; ++s) any |= e.hst[s].active != 0;
        if (!any) break;
        bool changed = false;
        if ((rc = regrow(e, changed))) return rc;
        if (changed) {
            drop_graph(e);
            hipLaunchKernelGGL(k_e_activate, dim3(nb(e.S)), dim3(256), 0, e.stream, e.S, e.st, target);
            SD_HIP(hipGetLastError());
        }
        if (use_graph) {
            if (!e.exec && (rc = capture(e))) return rc;
            SD_HIP(hipGraphLaunch(e.exec, e.stream));
        } else {
            if ((rc = enqueue_attempt(e))) return rc;
        }
        ++done;
        ++e.attempts;
    }
    if (done_attempts) *done_attempts = done;
    return read_state(e);
}

// Time of one launch configuration: speeds of every realization and one
// Newton solve at the current positions (step h) for each solver block.
int ensemble_autotune(Ensemble &e, std::string &report)
{
    hipEvent_t a, bq;
    SD_HIP(hipEventCreate(&a));
    SD_HIP(hipEventCreate(&bq));
    const long long keep = e.target;
    e.target = 1LL << 60;
    hipLaunchKernelGGL(k_e_activate, dim3(nb(e.S)), dim3(256), 0, e.stream, e.S, e.st, e.target);
    float best = 1e30f;
    int bestT = e.speedsBlock;
    char line[160];
    int rc;
    for (int T : {0, 64, 128, 256}) {
        if (T > 0 && (rc = build_blocks(e, T))) return rc;
        e.speedsBlock = T;
        float ms = 0;
        for (int r = 0; r < 3; ++r) {
            SD_HIP(hipEventRecord(a, e.stream));
            if ((rc = launch_speeds(e, ens::DISL, ens::VSPEED2, 0, 0))) return rc;
            SD_HIP(hipEventRecord(bq, e.stream));
            SD_HIP(hipEventSynchronize(bq));
            float t = 0;
            SD_HIP(hipEventElapsedTime(&t, a, bq));
            if (r > 0) ms += t;
        }
        ms /= 2;
        if (T == 0) snprintf(line, sizeof line, "speeds symmetric 64x64 tiles: %.3f ms\n", ms);
        else snprintf(line, sizeof line, "speeds one lane per target, block %4d: %.3f ms\n", T, ms);
        report += line;
        if (ms < best) {
            best = ms;
            bestT = T;
        }
    }
    if (bestT > 0 && (rc = build_blocks(e, bestT))) return rc;
    e.speedsBlock = bestT;
    // solver: a Newton system at the current positions (stage I's first
    // iteration), regrowing the CSC capacity if a count exceeds it
    for (int tries = 0;; ++tries) {
        hipLaunchKernelGGL(k_e_copy, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 0, e.x[ens::DISL],
                           e.x[ens::BIG]);
        if ((rc = launch_jacobian(e, ens::BIG, 1.0, true))) return rc;
        hipLaunchKernelGGL(k_e_residual, dim3(nb(e.NT)), dim3(256), 0, e.stream, e.NT, e.sys, e.st, 1.0,
                           e.x[ens::BIG], e.x[ens::DISL], e.v[ens::VINIT], e.v[ens::VINIT], e.dvec, e.g);
        if ((rc = read_state(e))) return rc;
        bool changed = false;
        if ((rc = regrow(e, changed))) return rc;
        if (!changed) break;
        SD_CHECK(tries < 4, SDDDST_ERR_STATE, "autotune: Jacobian capacity did not settle");
        hipLaunchKernelGGL(k_e_activate, dim3(nb(e.S)), dim3(256), 0, e.stream, e.S, e.st, e.target);
    }
    best = 1e30f;
    int bestB = e.solverBlock, bestL = e.spmvLanes;
    for (int B : {256, 512, 1024})
    for (int L : {4, 8, 16, 64}) {
        e.solverBlock = B;
        e.spmvLanes = L;
        float ms = 0;
        for (int r = 0; r < 3; ++r) {
            SD_HIP(hipEventRecord(a, e.stream));
            if ((rc = launch_solve(e))) return rc;
            SD_HIP(hipEventRecord(bq, e.stream));
            SD_HIP(hipEventSynchronize(bq));
            float t = 0;
            SD_HIP(hipEventElapsedTime(&t, a, bq));
            if (r > 0) ms += t;
        }
        ms /= 2;
        snprintf(line, sizeof line, "solver block %4d, %2d lanes per row: %.3f ms\n", B, L, ms);
        report += line;
        if (ms < best) {
            best = ms;
            bestB = B;
            bestL = L;
        }
    }
    e.solverBlock = bestB;
    e.spmvLanes = bestL;
    snprintf(line, sizeof line, "chosen: speeds %s %d, solver block %d, %d lanes per row\n",
             bestT == 0 ? "tiles" : "block", bestT, bestB, bestL);
    report += line;
    // the solves above only touched scratch; restore the counters
    if ((rc = read_state(e))) return rc;
    for (int s = 0; s < e.S; ++s) {
        e.hst[s].iters = 0;
        e.hst[s].solves = 0;
        e.hst[s].lu_fallbacks = 0;
        e.hst[s].solve_fail = 0;
    }
    SD_HIP(hipMemcpyAsync(e.st, e.hst, sizeof(SysState) * e.S, hipMemcpyHostToDevice, e.stream));
    e.target = keep;
    drop_graph(e);
    SD_HIP(hipStreamSynchronize(e.stream));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(bq);
    return SDDDST_OK;
}

int ensemble_state(Ensemble &e, double *t, double *h, int64_t *steps, int64_t *failed)
{
    int rc = read_state(e);
    if (rc) return rc;
    for (int s = 0; s < e.S; ++s) {
        if (t) t[s] = e.hst[s].t;
        if (h) h[s] = e.hst[s].h;
        if (steps) steps[s] = e.hst[s].steps;
        if (failed) failed[s] = e.hst[s].failed;
    }
    return SDDDST_OK;
}

int ensemble_positions(Ensemble &e, double *x)
{
    SD_HIP(hipMemcpyAsync(x, e.x[ens::DISL], sizeof(double) * e.NT, hipMemcpyDeviceToHost, e.stream));
    SD_HIP(hipStreamSynchronize(e.stream));
    return SDDDST_OK;
}

int ensemble_stats(Ensemble &e, int64_t *attempts, int64_t *iters, int64_t *solves, int64_t *lu, int64_t *regrows,
                   int *speeds_block, int *solver_block, int *spmv_lanes)
{
    int rc = read_state(e);
    if (rc) return rc;
    long long it = 0, so = 0, l = 0;
    for (int s = 0; s < e.S; ++s) {
        it += e.hst[s].iters;
        so += e.hst[s].solves;
        l += e.hst[s].lu_fallbacks + e.hst[s].solve_fail;
    }
    if (attempts) *attempts = e.attempts;
    if (iters) *iters = it;
    if (solves) *solves = so;
    if (lu) *lu = l;
    if (regrows) *regrows = e.regrows;
    if (speeds_block) *speeds_block = e.speedsBlock;
    if (solver_block) *solver_block = e.solverBlock;
    if (spmv_lanes) *spmv_lanes = e.spmvLanes;
    return SDDDST_OK;
}

}  // namespace sdamd

// ---------------------------------------------------------------------------
// C ABI

struct sdddst_ensemble {
    sdamd::Ensemble *e;
};

extern "C" {

int sdddst_ensemble_create(int num_systems, const int *n, const double *x, const double *y, const double *b,
                           const sdddst_params *params, double sigma, double initial_step, sdddst_ensemble **out)
{
    SD_CHECK(out, SDDDST_ERR_ARG, "out is NULL");
    sdamd::Ensemble *e = nullptr;
    int rc = sdamd::ensemble_create(num_systems, n, x, y, b, params, sigma, initial_step, &e);
    if (rc) return rc;
    *out = new sdddst_ensemble{e};
    return SDDDST_OK;
}

int sdddst_ensemble_destroy(sdddst_ensemble *ens)
{
    if (!ens) return SDDDST_OK;
    sdamd::ensemble_free(ens->e);
    delete ens;
    return SDDDST_OK;
}

int sdddst_ensemble_set_tiles(sdddst_ensemble *ens, int speeds_block, int solver_block, int spmv_lanes)
{
    SD_CHECK(ens && ens->e, SDDDST_ERR_ARG, "ensemble is NULL");
    return sdamd::ensemble_set_tiles(*ens->e, speeds_block, solver_block, spmv_lanes);
}

int sdddst_ensemble_autotune(sdddst_ensemble *ens, char *report, int report_len)
{
    SD_CHECK(ens && ens->e, SDDDST_ERR_ARG, "ensemble is NULL");
    std::string r;
    int rc = sdamd::ensemble_autotune(*ens->e, r);
    if (report && report_len > 0) {
        strncpy(report, r.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return rc;
}

int sdddst_ensemble_run(sdddst_ensemble *ens, int64_t target_steps, int64_t max_attempts, int use_graph,
                        int64_t *attempts)
{
    SD_CHECK(ens && ens->e, SDDDST_ERR_ARG, "ensemble is NULL");
    SD_CHECK(target_steps >= 0 && max_attempts >= 0, SDDDST_ERR_ARG, "negative step counts");
    long long done = 0;
    int rc = sdamd::ensemble_run(*ens->e, target_steps, max_attempts, use_graph != 0, &done);
    if (attempts) *attempts = done;
    return rc;
}

int sdddst_ensemble_state(sdddst_ensemble *ens, double *t, double *h, int64_t *steps, int64_t *failed)
{
    SD_CHECK(ens && ens->e, SDDDST_ERR_ARG, "ensemble is NULL");
    return sdamd::ensemble_state(*ens->e, t, h, steps, failed);
}

int sdddst_ensemble_positions(sdddst_ensemble *ens, double *x)
{
    SD_CHECK(ens && ens->e && x, SDDDST_ERR_ARG, "bad arguments");
    return sdamd::ensemble_positions(*ens->e, x);
}

int sdddst_ensemble_stats(sdddst_ensemble *ens, int64_t *attempts, int64_t *solver_iters, int64_t *solves,
                          int64_t *lu_fallbacks, int64_t *regrows, int *speeds_block, int *solver_block,
                          int *spmv_lanes)
{
    SD_CHECK(ens && ens->e, SDDDST_ERR_ARG, "ensemble is NULL");
    return sdamd::ensemble_stats(*ens->e, attempts, solver_iters, solves, lu_fallbacks, regrows, speeds_block,
                                 solver_block, spmv_lanes);
}

}  // extern "C"
