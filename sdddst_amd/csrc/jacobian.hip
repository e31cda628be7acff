// K2: cutoff-limited Jacobian assembly (Simulation::calculateJacobian,
// simulation.cpp:236-384) into the reference's CSC arrays Ap[N+1], Ai[nnz],
// Ax[nnz] with the reference's exact sparsity pattern.
//
// Pattern.  Column j holds, with rows ascending: the rows i < j whose pair
// passed the predicate and whose value is nonzero (simulation.cpp:260-268),
// the diagonal j, and the rows i > j that pass the predicate
// (simulation.cpp:327-345).  The predicate of the pair {lo < hi} is
//     (sqrt(dx*dx + dy*dy) - rc)^2 < 36.8 * rc^2,
// dx = wrap(x_hi - x_lo), dy = wrap(y_hi - y_lo), every operation rounded on
// its own (no contraction), pow(., 2) == ., * . as an optimised build folds
// it.  The pattern is therefore symmetric except when a lower value is
// exactly zero; that case is detected and compacted afterwards.
//
// Stages (all on the context stream):
//   count  one wavefront per column; with cells at least half the
//          pre-filter radius wide, the candidates are the 5 x 5 neighbouring
//          cells' dislocations (cell-ordered copies of their positions), else all
//          rows 64 at a time; a cheap r^2 pre-filter precedes the square
//          root; the column's membership bitmask is stored, so the predicate
//          runs once per ordered pair;
//   scan   one workgroup, exclusive prefix sum into Ap (int64 checked);
//   fill   one wavefront per column, from the bitmask: prefix popcounts give
//          each row its slot, rows land in ascending order, the diagonal
//          position and each mask word's first position are recorded;
//   values one wavefront per column over its strictly-lower rows: tau_x and
//          the taper once per unique pair, kept with the position of the
//          mirror entry (row j of column i: the word start plus a popcount);
//   rawsum one wavefront per column: the unscaled values h*b_i*b_j*tau_x*
//          taper in storage order (upper entries through their mirrors),
//          stored coalesced, and their ordered sum
//          (simulation.cpp:349-374) giving the raw diagonal and d_j;
//   scale  one wavefront per column: A = row-scaled values (+1 on the
//          diagonal, simulation.cpp:376-383) stored row-major (AxRow: row j
//          of A on the index arrays of column j), which the solvers read, and
//          the pair partner of row j; the column-major Ax the reference
//          returns is formed from the same factors on request (ensure_ax:
//          downloads, the dense solver).  Stage II of a step assembles at the
//          same positions as stage I with h/2: its unscaled values and sums
//          are stage I's times 0.5 exactly, so only d_j and scale rerun.
//
// Algorithmic HBM traffic of one assembly: 12*nnz + 40*N + 4 bytes.
#include <hip/hip_runtime.h>

#include <cstring>

#include <vector>

#include "context.h"
#include "field.h"
#include "jac_core.h"

namespace sdamd {

namespace {

using jac::CutParams;

__global__ __launch_bounds__(256) void k_jac_count(int n, int words, const double *__restrict__ x,
                                                   const double *__restrict__ y, CutParams cp,
                                                   int *__restrict__ cnt, unsigned long long *__restrict__ mask,
                                                   unsigned long long *__restrict__ cmask, int *__restrict__ ccnt)
{
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    jac::count_column(j, n, words, x, y, cp, cnt + j, mask + (size_t)j * words,
                      cmask ? cmask + (size_t)j * words : nullptr, cmask ? ccnt + j : nullptr);
}

// Cell binning for the count: the unit cell is cut into G x G periodic cells
// at least the predicate's pre-filter radius wide (with a 1e-6 margin, so no
// rounding of a cell index can separate a qualifying pair by two cells);
// every pair that can pass the predicate lies in neighbouring cells.
__device__ __forceinline__ int cell_of(double x, double y, int G)
{
    const double xw = sdfield::normalize(x), yw = sdfield::normalize(y);
    int cx = (int)((xw + 0.5) * G), cy = (int)((yw + 0.5) * G);
    cx = cx < 0 ? 0 : (cx >= G ? G - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= G ? G - 1 : cy);
    return cy * G + cx;
}

// Cell of each dislocation and the cells' counts: each workgroup counts its
// kCellPer * 256 dislocations in LDS and adds one total per occupied cell.
constexpr int kCellPer = 4;
constexpr int kCellMaxG = 64;   // G * G counts (and bases) fit in LDS

__global__ __launch_bounds__(256) void k_cell_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                    int G, int *__restrict__ cell, int *__restrict__ ccount)
{
    extern __shared__ int hist[];
    const int cells = G * G, t = (int)threadIdx.x;
    for (int k = t; k < cells; k += 256) hist[k] = 0;
    __syncthreads();
    const int i0 = (int)blockIdx.x * 256 * kCellPer + t;
#pragma unroll
    for (int e = 0; e < kCellPer; ++e) {
        const int i = i0 + e * 256;
        if (i < n) {
            const int c = cell_of(x[i], y[i], G);
            cell[i] = c;
            atomicAdd(&hist[c], 1);
        }
    }
    __syncthreads();
    for (int k = t; k < cells; k += 256)
        if (hist[k]) atomicAdd(&ccount[k], hist[k]);
}

// cstart = exclusive scan of ccount (G*G cells, one workgroup); the fill
// cursor starts at cstart, and ccount is left zero for the next count.
__global__ __launch_bounds__(1024) void k_cell_scan(int cells, int *__restrict__ ccount,
                                                    int *__restrict__ cstart, int *__restrict__ cursor)
{
    __shared__ int part[1024];
    const int t = (int)threadIdx.x;
    const int per = (cells + 1023) / 1024;
    const int s0 = t * per, s1 = min(cells, s0 + per);
    int acc = 0;
    for (int k = s0; k < s1; ++k) acc += ccount[k];
    part[t] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - acc;
    for (int k = s0; k < s1; ++k) {
        cstart[k] = run;
        cursor[k] = run;
        run += ccount[k];
        ccount[k] = 0;
    }
    if (t == 1023) cstart[cells] = part[1023];
}

// Dislocations by cell, with copies of their positions in the same order
// (the order within a cell is arbitrary: the count only sets bits).  Each
// workgroup ranks its dislocations per cell in LDS and reserves one range
// per occupied cell.
__global__ __launch_bounds__(256) void k_cell_scatter(int n, int G, const int *__restrict__ cell,
                                                      const double *__restrict__ x, const double *__restrict__ y,
                                                      int *__restrict__ cursor, int *__restrict__ sorted,
                                                      double *__restrict__ sx, double *__restrict__ sy)
{
    extern __shared__ int lds[];
    const int cells = G * G, t = (int)threadIdx.x;
    int *hist = lds, *base = lds + cells;
    for (int k = t; k < cells; k += 256) hist[k] = 0;
    __syncthreads();
    const int i0 = (int)blockIdx.x * 256 * kCellPer + t;
    int c[kCellPer], r[kCellPer];
#pragma unroll
    for (int e = 0; e < kCellPer; ++e) {
        const int i = i0 + e * 256;
        c[e] = i < n ? cell[i] : -1;
        r[e] = c[e] >= 0 ? atomicAdd(&hist[c[e]], 1) : 0;
    }
    __syncthreads();
    for (int k = t; k < cells; k += 256)
        if (hist[k]) base[k] = atomicAdd(&cursor[k], hist[k]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kCellPer; ++e) {
        const int i = i0 + e * 256;
        if (c[e] >= 0) {
            const int p = base[c[e]] + r[e];
            sorted[p] = i;
            sx[p] = x[i];
            sy[p] = y[i];
        }
    }
}

// Count of column j from the neighbouring cells: the same predicate as
// count_column, bits gathered in an LDS copy of the column's bitmask (the
// order of the candidates does not matter), then written out whole; with
// cmask, the within-cutoff bitmask of count_column the same way (and its
// count into ccnt).
constexpr int kCellMaxWords = 1024;
// cells at least 1 / kCellReach of the pre-filter radius wide: a column's
// candidates are the (2 kCellReach + 1)^2 cells around its own (25 cells of
// width rpre / 2 cover 6.25 rpre^2, against 9 rpre^2 for 3 x 3 cells of
// width rpre)
constexpr int kCellReach = 2;
__global__ __launch_bounds__(256) void k_jac_count_cells(int n, int words, const double *__restrict__ x,
                                                         const double *__restrict__ y, CutParams cp, int G,
                                                         const int *__restrict__ cell,
                                                         const int *__restrict__ cstart,
                                                         const int *__restrict__ sorted,
                                                         const double *__restrict__ sx,
                                                         const double *__restrict__ sy, int *__restrict__ cnt,
                                                         unsigned long long *__restrict__ mask,
                                                         unsigned long long *__restrict__ cmask,
                                                         int *__restrict__ ccnt)
{
    extern __shared__ unsigned long long lm[];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int j = (int)(blockIdx.x * 4 + wave);
    unsigned long long *m = lm + (size_t)wave * words;
    unsigned long long *cm = lm + (size_t)(4 + wave) * words;   // used with cmask only
    for (int w = lane; w < words; w += 64) m[w] = 0ull;
    if (cmask)
        for (int w = lane; w < words; w += 64) cm[w] = 0ull;
    __syncthreads();
    if (j < n) {
        const double xj = x[j], yj = y[j];
        const int c = cell[j];
        const int cx = c % G, cy = c / G;
        constexpr int R = kCellReach;
        for (int oy = -R; oy <= R; ++oy) {
            const int yy = (cy + oy + G) % G;
            // the 2R + 1 cells of this row: one contiguous range unless the
            // row wraps around the cell edge
            const bool inner = cx >= R && cx <= G - 1 - R;
            for (int ox = inner ? 0 : -R; ox <= (inner ? 0 : R); ++ox) {
                const int q = yy * G + (cx + ox + G) % G;
                const int k0 = inner ? cstart[q - R] : cstart[q], k1 = inner ? cstart[q + R + 1] : cstart[q + 1];
                for (int k = k0 + lane; k < k1; k += 64) {
                    const int i = sorted[k];
                    const double xi = sx[k], yi = sy[k];
                    bool in, core = false;
                    if (i == j) {
                        in = true;
                    } else {
                        const double dx = sdfield::normalize(i > j ? xi - xj : xj - xi);
                        const double dy = sdfield::normalize(i > j ? yi - yj : yj - yi);
                        double r2;
                        in = jac::pair_pred(dx, dy, cp, r2);
                        core = !(r2 > cp.rc2);
                    }
                    if (in) atomicOr(&m[i >> 6], 1ull << (i & 63));
                    if (core && cmask) atomicOr(&cm[i >> 6], 1ull << (i & 63));
                }
            }
        }
    }
    __syncthreads();
    if (j >= n) return;
    int c = 0;
    unsigned long long *mj = mask + (size_t)j * words;
    for (int w = lane; w < words; w += 64) {
        const unsigned long long v = m[w];
        mj[w] = v;
        c += __popcll(v);
    }
    if (cmask) {
        unsigned long long *cj = cmask + (size_t)j * words;
        int cc = 0;
        for (int w = lane; w < words; w += 64) {
            const unsigned long long v = cm[w];
            cj[w] = v;
            cc += __popcll(v);
        }
        for (int off = 32; off > 0; off >>= 1) cc += __shfl_down(cc, off, 64);
        if (lane == 0) ccnt[j] = cc;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if (lane == 0) cnt[j] = c;
}

// Exclusive scan of n counts into Ap[0..n] (and the total) by one workgroup
// of 1024 lanes over tiles of 16384 counts: each lane sums 16 consecutive
// counts (four 16-byte loads), each wave scans its 64 lane sums with
// shuffles, the first wave scans the 16 wave sums, and a 64-bit carry joins
// the tiles (three barriers per tile).
__global__ __launch_bounds__(1024) void k_scan(int n, const int *__restrict__ cnt, int *__restrict__ Ap,
                                               long long *__restrict__ total)
{
    constexpr int kPer = 16, kTileN = 1024 * kPer;
    __shared__ long long wsum[16];
    const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
    const bool vec = ((reinterpret_cast<uintptr_t>(cnt) | reinterpret_cast<uintptr_t>(Ap)) & 15) == 0;
    long long carry = 0;
    for (int base = 0; base < n; base += kTileN) {
        const int q0 = base + t * kPer;
        const bool full = vec && q0 + kPer <= n;
        int v[kPer];
        if (full) {
#pragma unroll
            for (int u = 0; u < kPer / 4; ++u) {
                const int4 a = reinterpret_cast<const int4 *>(cnt + q0)[u];
                v[4 * u] = a.x;
                v[4 * u + 1] = a.y;
                v[4 * u + 2] = a.z;
                v[4 * u + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int u = 0; u < kPer; ++u) v[u] = q0 + u < n ? cnt[q0 + u] : 0;
        }
        long long acc = 0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) acc += v[u];
        long long incl = acc;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            long long sw = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const long long o = __shfl_up(sw, off, 64);
                if (lane >= off) sw += o;
            }
            if (lane < 16) wsum[lane] = sw;
        }
        __syncthreads();
        long long run = carry + (w > 0 ? wsum[w - 1] : 0) + incl - acc;   // exclusive prefix of this lane's counts
        int o[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            o[u] = (int)run;
            run += v[u];
        }
        if (full) {
#pragma unroll
            for (int u = 0; u < kPer / 4; ++u)
                reinterpret_cast<int4 *>(Ap + q0)[u] = make_int4(o[4 * u], o[4 * u + 1], o[4 * u + 2], o[4 * u + 3]);
        } else {
#pragma unroll
            for (int u = 0; u < kPer; ++u)
                if (q0 + u < n) Ap[q0 + u] = o[u];
        }
        carry += wsum[15];
        __syncthreads();   // wsum is rewritten by the next tile
    }
    if (t == 0) {
        *total = carry;
        Ap[n] = (int)carry;
    }
}

// The kernels after the entry count run only while it fits the storage the
// host allocated (total <= cap); the host reads the count with the
// assembly's one wait and, on overflow, grows the storage and assembles
// again.
__device__ __forceinline__ bool over_capacity(const long long *total, long long cap) { return *total > cap; }

__global__ __launch_bounds__(256) void k_jac_fill(int n, int words, const unsigned long long *__restrict__ mask,
                                                  const int *__restrict__ Ap, int *__restrict__ Ai,
                                                  int *__restrict__ diag, int *__restrict__ wstart,
                                                  const long long *__restrict__ total, long long cap)
{
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n || over_capacity(total, cap)) return;
    jac::fill_column(j, words, mask + (size_t)j * words, Ap[j], Ai, diag + j, wstart + (size_t)j * words);
}

// tau_x and the taper of every strictly-lower entry (row i > column j), the
// position of its mirror (row j in column i), and the unsigned value for step
// size h at both positions.
template <int Mode, bool Aux>
__global__ __launch_bounds__(256) void k_jac_values(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                    CutParams cp, const int *__restrict__ Ap,
                                                    const int *__restrict__ Ai, const int *__restrict__ diag,
                                                    double *__restrict__ tau, double *__restrict__ mult,
                                                    int *__restrict__ mirror, double *__restrict__ xsnap,
                                                    int words, const unsigned long long *__restrict__ mask,
                                                    const int *__restrict__ wstart, const double *__restrict__ b,
                                                    double h, double *__restrict__ raw, int *__restrict__ flags,
                                                    const long long *__restrict__ total, long long cap)
{
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n || over_capacity(total, cap)) return;
    jac::values_column<Mode, Aux>(j, x, y, cp, Ap, Ai, diag, tau, mult, mirror, xsnap, words, mask, wstart, b, h,
                                   raw, flags);
}

// simulation.cpp:269-324: the point-defect sum tmp of column j (defects in
// ascending order, tmp -= b_j * dF/dx * taper), h-independent.
__global__ void k_jac_pd(int n, int pc, const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ b, const double *__restrict__ px,
                         const double *__restrict__ py, double A, double K, CutParams cp,
                         double *__restrict__ pdtmp)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    const double xj = x[j], yj = y[j], bj = b[j];
    double tmp = 0;
    for (int l = 0; l < pc; ++l) {
        const double dx = sdfield::normalize(xj - px[l]);
        const double dy = sdfield::normalize(yj - py[l]);
        double r2;
        if (jac::pair_pred(dx, dy, cp, r2)) {
            double multiplier = 1;
            if (r2 > cp.rc2) {
                const double s = sdlibm::sqrt(r2) - cp.rc;
                multiplier = sdlibm::exp(-(s * s) * cp.inv_rc2);
            }
            tmp -= bj * sdfield::pd_force_dx(dx, dy, A, K) * multiplier;
        }
    }
    pdtmp[j] = tmp;
}

__global__ __launch_bounds__(256) void k_jac_raw(int n, const double *__restrict__ b, double h,
                                                 const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                 const int *__restrict__ diag, const double *__restrict__ tau,
                                                 const double *__restrict__ mult, const int *__restrict__ mirror,
                                                 const double *__restrict__ pdtmp, double *__restrict__ Ax,
                                                 int *__restrict__ flags)
{
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    jac::raw_column(j, b, h, Ap, Ai, diag, tau, mult, mirror, pdtmp ? pdtmp[j] : 0.0, Ax, flags);
}

__global__ __launch_bounds__(256) void k_jac_diag(int n, const int *__restrict__ Ap, const int *__restrict__ diag,
                                                  double *__restrict__ Ax, double *__restrict__ dvec)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    jac::diag_column(j, Ap, diag, Ax, dvec);
}

__global__ __launch_bounds__(256) void k_jac_scale(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                   const int *__restrict__ diag, const double *__restrict__ dvec,
                                                   double *__restrict__ Ax, double *__restrict__ AxRow)
{
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    jac::scale_column(j, Ap, Ai, diag, dvec, Ax, AxRow);
}

// Pass 2 of the reference (simulation.cpp:349-374) over the unsigned values
// already in storage order (Araw, written at both positions of every pair by
// the values pass or by k_jac_raw): one wavefront per column j stages the
// column in LDS with coalesced loads, the diagonal taking -pd * h, and lane 0
// adds it in storage order, giving s = -sum, the raw diagonal, and d_j.  The
// next chunk's loads are in flight while lane 0 adds the current one.
constexpr int kSumChunk = 512;
constexpr int kSumPer = kSumChunk / 64;
__global__ __launch_bounds__(256) void k_jac_rawsum(int n, double h, const int *__restrict__ Ap,
                                                    const int *__restrict__ diag, const double *__restrict__ pdtmp,
                                                    double *__restrict__ Araw, double *__restrict__ sraw,
                                                    double *__restrict__ dvec, double2 *__restrict__ colinfo,
                                                    const long long *__restrict__ total, long long cap,
                                                    int *__restrict__ strong)
{
    __shared__ double buf[4][kSumChunk];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int j = (int)(blockIdx.x * 4 + wave);
    if (j >= n || over_capacity(total, cap)) return;
    const int beg = Ap[j], end = Ap[j + 1], d = diag[j];
    const double pdh = -(pdtmp ? pdtmp[j] : 0.0) * h;
    double *lb = buf[wave];
    double subSum = 0;
    double nx[kSumPer];
#pragma unroll
    for (int u = 0; u < kSumPer; ++u) {
        const int k = beg + lane + 64 * u;
        nx[u] = k < end ? (k == d ? pdh : Araw[k]) : 0.0;
    }
    for (int c0 = beg; c0 < end; c0 += kSumChunk) {
        const int c1 = min(end, c0 + kSumChunk);
#pragma unroll
        for (int u = 0; u < kSumPer; ++u) lb[lane + 64 * u] = nx[u];
        // the wavefront's LDS stores are visible to its own later loads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (c1 < end) {
#pragma unroll
            for (int u = 0; u < kSumPer; ++u) {
                const int k = c1 + lane + 64 * u;
                nx[u] = k < end ? (k == d ? pdh : Araw[k]) : 0.0;
            }
        }
        if (lane == 0)
            for (int t = 0; t < c1 - c0; ++t) subSum += lb[t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane != 0) return;
    subSum *= -1.;
    Araw[d] = subSum;
    sraw[j] = subSum;
    double dv;
    if (subSum > 0) {
        double t = 1. / subSum;
        t += 1.;
        t *= t;
        dv = 1. / t;
    } else {
        dv = 0.;
    }
    dvec[j] = dv;
    const double Dj = (1.0 + dv) * 0.5;
    colinfo[j] = make_double2(Dj, __builtin_fabs(subSum * Dj + 1.0));
    // a diagonal far from 1 marks a matrix far from the identity (a step
    // size large for the configuration), counted for the preconditioner
    // choice
    if (__builtin_fabs(subSum * Dj) > kStrongDiag) atomicAdd(strong, 1);
}

// d_j for the column sums s * factor (factor a power of two: the unscaled
// values and their ordered sums scale exactly).
__global__ void k_jac_dvec(int n, const double *__restrict__ sraw, double factor, double *__restrict__ dvec,
                           double2 *__restrict__ colinfo)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    const double s = sraw[j] * factor;
    double dv;
    if (s > 0) {
        double t = 1. / s;
        t += 1.;
        t *= t;
        dv = 1. / t;
    } else {
        dv = 0.;
    }
    dvec[j] = dv;
    const double Dj = (1.0 + dv) * 0.5;
    colinfo[j] = make_double2(Dj, __builtin_fabs(s * Dj + 1.0));
}

// Fused pass 2 (simulation.cpp:376-383) for column j with the unscaled values
// Araw * factor: Ax = raw * (1 + d_i) / 2 and AxRow = raw * (1 + d_j) / 2,
// +1 on the diagonal; and the pair partner of row j from the same values
// (the rule of partner_row in precond_core.h: the j' maximising
// |A_jj' A_j'j| / (|A_jj| |A_j'j'|), ties to the smaller j', with |A_j'j'|
// formed as the scaling forms it).  The row scales and diagonal magnitudes
// come packed per column (one 16-byte gather per entry), and each lane
// issues the loads of four entries before their arithmetic, so the
// index -> gather latency chain is paid once per four entries.
constexpr int kScaleUnroll = 4;
__global__ __launch_bounds__(256) void k_jac_scale_partner(int n, const int *__restrict__ Ap,
                                                           const int *__restrict__ Ai,
                                                           const int *__restrict__ diag,
                                                           const double2 *__restrict__ colinfo,
                                                           const double *__restrict__ Araw, double factor,
                                                           double *__restrict__ Ax, double *__restrict__ AxRow,
                                                           int *__restrict__ partner, double *__restrict__ strength,
                                                           unsigned short *__restrict__ Ai16)
{
    const int lane = (int)(threadIdx.x & 63);
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    const double2 cj = colinfo[j];
    const double Dj = cj.x, ajj = cj.y;
    const int dpos = diag[j];
    const int end = Ap[j + 1];
    double best = -1.0;
    int bj = -1;
    for (int k0 = Ap[j] + lane; k0 < end; k0 += 64 * kScaleUnroll) {
        double raw[kScaleUnroll];
        int ii[kScaleUnroll];
#pragma unroll
        for (int u = 0; u < kScaleUnroll; ++u) {
            const int k = k0 + 64 * u;
            raw[u] = k < end ? Araw[k] : 0.0;
            ii[u] = k < end ? Ai[k] : j;
        }
        double2 ci[kScaleUnroll];
#pragma unroll
        for (int u = 0; u < kScaleUnroll; ++u) ci[u] = colinfo[ii[u]];
#pragma unroll
        for (int u = 0; u < kScaleUnroll; ++u) {
            const int k = k0 + 64 * u;
            if (k >= end) continue;
            const double r0 = raw[u] * factor;
            const int i = ii[u];
            double a = r0 * ci[u].x;
            double r = r0 * Dj;
            if (k == dpos) {
                a += 1.0;
                r += 1.0;
            }
            if (Ax) __builtin_nontemporal_store(a, &Ax[k]);
            AxRow[k] = r;
            if (Ai16) Ai16[k] = (unsigned short)i;
            if (i != j) {
                const double st = __builtin_fabs(r * a) / (ajj * ci[u].y);
                if (st > best || (st == best && i < bj)) {
                    best = st;
                    bj = i;
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_down(best, off, 64);
        const int oj = __shfl_down(bj, off, 64);
        if (ob > best || (ob == best && oj >= 0 && (bj < 0 || oj < bj))) {
            best = ob;
            bj = oj;
        }
    }
    if (lane == 0) {
        partner[j] = bj;
        strength[j] = best;
    }
}

// The CSC values A = (factor * unsigned) * row scale (+1 on the diagonal),
// formed exactly as the scaling pass forms them, for the callers that read
// the column-major values (downloads, the dense solver, training).
__global__ __launch_bounds__(256) void k_jac_ax(int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                                const int *__restrict__ diag, const double2 *__restrict__ colinfo,
                                                const double *__restrict__ Araw, double factor,
                                                double *__restrict__ Ax)
{
    const int lane = (int)(threadIdx.x & 63);
    const int j = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (j >= n) return;
    const int dpos = diag[j];
    const int end = Ap[j + 1];
    for (int k = Ap[j] + lane; k < end; k += 64) {
        const double r0 = Araw[k] * factor;
        double a = r0 * colinfo[Ai[k]].x;
        if (k == dpos) a += 1.0;
        Ax[k] = a;
    }
}

// the zero-value flag and the strong-diagonal count of the assembly flags
__global__ void k_flags_reset(int *f)
{
    if (threadIdx.x == 0) {
        f[0] = 0;
        f[4] = 0;
    }
}

}  // namespace

static CutParams cut_params(const Context &c)
{
    return jac::make_cut_params(c.cutoff, c.cutoff_sqr, c.one_per_cutoff_sqr);
}

int launch_scan(Context &c, int n, const int *cnt, int *out, long long *d_total)
{
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c.stream, n, cnt, out, d_total);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

// Rebuild without the upper entries whose mirrored value is exactly zero
// (getElement returns 0 for them, simulation.cpp:263).  Host path: it runs
// only when such a value occurred.
static int compact_zero_upper(Context &c)
{
    const int n = c.n;
    std::vector<int> Ap(n + 1), Ai(c.nnz), diag(n);
    std::vector<double> Ax(c.nnz);
    SD_HIP(hipMemcpyAsync(Ap.data(), c.Ap, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipMemcpyAsync(Ai.data(), c.Ai, sizeof(int) * c.nnz, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipMemcpyAsync(Ax.data(), c.Ax, sizeof(double) * c.nnz, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> nAp(n + 1), nAi, ndiag(n);
    std::vector<double> nAx;
    nAi.reserve(c.nnz);
    nAx.reserve(c.nnz);
    nAp[0] = 0;
    for (int j = 0; j < n; ++j) {
        for (int k = Ap[j]; k < Ap[j + 1]; ++k) {
            const int i = Ai[k];
            if (i < j && Ax[k] == 0.0) continue;
            if (i == j) ndiag[j] = (int)nAi.size();
            nAi.push_back(i);
            nAx.push_back(Ax[k]);
        }
        nAp[j + 1] = (int)nAi.size();
    }
    c.nnz = (int64_t)nAi.size();
    c.ai16_valid = false;   // the kept entries moved
    SD_HIP(hipMemcpyAsync(c.Ap, nAp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipMemcpyAsync(c.Ai, nAi.data(), sizeof(int) * c.nnz, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipMemcpyAsync(c.Ax, nAx.data(), sizeof(double) * c.nnz, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipMemcpyAsync(c.diag, ndiag.data(), sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

static int ensure_capacity(Context &c, int64_t need)
{
    if (need <= c.cap) return SDDDST_OK;
    int64_t ncap = need + need / 8 + 1024;
    void **bufs[] = {(void **)&c.Ai,      (void **)&c.Ax,   (void **)&c.AxRow, (void **)&c.jtau, (void **)&c.jmult,
                     (void **)&c.jmirror, (void **)&c.Araw, (void **)&c.Ai16};
    for (void **p : bufs) {
        if (*p) SD_HIP(hipFree(*p));
        *p = nullptr;
    }
    c.cap = 0;
    c.raw_valid = false;
    c.ai16_valid = false;
    c.araw_h = __builtin_nan("");
    const size_t sizes[] = {sizeof(int), sizeof(double), sizeof(double), sizeof(double),
                            sizeof(double), sizeof(int), sizeof(double), sizeof(unsigned short)};
    // 16-bit row indices only where every index fits
    const int nbufs = c.n <= 65536 ? 8 : 7;
    for (int k = 0; k < nbufs; ++k) {
        if (hipMalloc(bufs[k], sizes[k] * ncap) != hipSuccess) {
            *bufs[k] = nullptr;
            set_error("device allocation of %lld Jacobian entries failed", (long long)ncap);
            return SDDDST_ERR_NOMEM;
        }
    }
    c.cap = ncap;
    return SDDDST_OK;
}

// Pattern, tau_x, taper and mirror positions for the positions in slot, and
// the off-diagonal unsigned values for step size h (with the zero flag).
static int build_pattern(Context &c, int slot, double h)
{
    const int n = c.n;
    const CutParams cp = cut_params(c);
    const double *x = c.x[slot];
    const int wblocks = (n + 3) / 4;
    const int words = (n + 63) / 64;
    if (!c.jmask && hipMalloc(&c.jmask, sizeof(unsigned long long) * (size_t)n * words) != hipSuccess) {
        c.jmask = nullptr;
        set_error("device allocation of the %d x %d membership bitmask failed", n, words);
        return SDDDST_ERR_NOMEM;
    }
    if (!c.jwstart && hipMalloc(&c.jwstart, sizeof(int) * (size_t)n * words) != hipSuccess) {
        c.jwstart = nullptr;
        set_error("device allocation of the %d x %d mask word starts failed", n, words);
        return SDDDST_ERR_NOMEM;
    }
    // the within-cutoff bitmask, for the learned preconditioner's core graph
    const bool want_core = c.precond == SDDDST_PRECOND_LEARNED;
    if (want_core && !c.jcmask && hipMalloc(&c.jcmask, sizeof(unsigned long long) * (size_t)n * words) != hipSuccess) {
        c.jcmask = nullptr;
        set_error("device allocation of the %d x %d within-cutoff bitmask failed", n, words);
        return SDDDST_ERR_NOMEM;
    }
    if (want_core && !c.jccnt) SD_HIP(hipMalloc(&c.jccnt, sizeof(int) * n));
    unsigned long long *cmask = want_core ? c.jcmask : nullptr;
    // cells at least rpre / kCellReach wide; the neighbourhoods need G >= 2 kCellReach + 1
    const double rpre = sqrt(cp.prefilter);
    // (fewer, wider cells are as valid: at most kCellMaxG per side)
    const int G = rpre > 0 && rpre < 1.0 / 3.0
                      ? min(kCellMaxG, (int)floor((double)kCellReach / (rpre * (1.0 + 1e-6))))
                      : 0;
    if (G >= 2 * kCellReach + 2 && words <= kCellMaxWords) {
        const int cells = G * G;
        if (cells > c.cell_cap) {
            for (int **q : {&c.ccount, &c.cstart, &c.ccursor}) {
                if (*q) SD_HIP(hipFree(*q));
                *q = nullptr;
                SD_HIP(hipMalloc(q, sizeof(int) * (cells + 1)));
            }
            // the counts start at zero; each scan leaves them zero again
            SD_HIP(hipMemsetAsync(c.ccount, 0, sizeof(int) * (cells + 1), c.stream));
            c.cell_cap = cells;
        }
        if (!c.cell) SD_HIP(hipMalloc(&c.cell, sizeof(int) * n));
        if (!c.csorted) SD_HIP(hipMalloc(&c.csorted, sizeof(int) * n));
        if (!c.csx) SD_HIP(hipMalloc(&c.csx, sizeof(double) * n));
        if (!c.csy) SD_HIP(hipMalloc(&c.csy, sizeof(double) * n));
        const int cblocks = (n + 256 * kCellPer - 1) / (256 * kCellPer);
        hipLaunchKernelGGL(k_cell_count, dim3(cblocks), dim3(256), sizeof(int) * cells, c.stream, n, x, c.y, G,
                           c.cell, c.ccount);
        hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, c.stream, cells, c.ccount, c.cstart, c.ccursor);
        hipLaunchKernelGGL(k_cell_scatter, dim3(cblocks), dim3(256), 2 * sizeof(int) * cells, c.stream, n, G, c.cell,
                           x, c.y, c.ccursor, c.csorted, c.csx, c.csy);
        hipLaunchKernelGGL(k_jac_count_cells, dim3(wblocks), dim3(256),
                           sizeof(unsigned long long) * (cmask ? 8 : 4) * words, c.stream, n, words, x, c.y, cp, G,
                           c.cell, c.cstart, c.csorted, c.csx, c.csy, c.colcount, c.jmask, cmask, c.jccnt);
    } else {
        hipLaunchKernelGGL(k_jac_count, dim3(wblocks), dim3(256), 0, c.stream, n, words, x, c.y, cp, c.colcount,
                           c.jmask, cmask, c.jccnt);
    }
    c.jcmask_valid = want_core;
    SD_HIP(hipGetLastError());
    long long *d_total = reinterpret_cast<long long *>(c.d_flags + 2);
    int rc = launch_scan(c, n, c.colcount, c.Ap, d_total);
    if (rc) return rc;
    if (c.cap == 0) {
        // the first assembly sizes the storage; later ones run against the
        // capacity and the count is read with the assembly's one wait
        long long total = 0;
        if ((rc = host_wait(c, d_total, 1, &total))) return rc;
        SD_CHECK(total < (1LL << 31) - 1, SDDDST_ERR_ARG, "nnz %lld exceeds the int32 index range", total);
        if ((rc = ensure_capacity(c, total))) return rc;
    }
    hipLaunchKernelGGL(k_jac_fill, dim3(wblocks), dim3(256), 0, c.stream, n, words, c.jmask, c.Ap, c.Ai, c.diag,
                       c.jwstart, d_total, (long long)c.cap);
    SD_HIP(hipGetLastError());
    c.ai16_valid = false;   // the scaling pass copies the new row indices
    hipLaunchKernelGGL(k_flags_reset, dim3(1), dim3(64), 0, c.stream, c.d_flags);
    // tau_x, the taper and the mirror positions are kept only when a later
    // call needs them (ensure_aux): the step never does, and the three arrays
    // were a fifth of the values pass's stores
    if (c.p.field_mode == SDDDST_FIELD_FAST)
        hipLaunchKernelGGL((k_jac_values<SDDDST_FIELD_FAST, false>), dim3(wblocks), dim3(256), 0, c.stream, n, x, c.y,
                           cp, c.Ap, c.Ai, c.diag, c.jtau, c.jmult, c.jmirror, c.jx, words, c.jmask, c.jwstart,
                           c.b, h, c.Araw, c.d_flags, d_total, (long long)c.cap);
    else
        hipLaunchKernelGGL((k_jac_values<SDDDST_FIELD_EXACT, false>), dim3(wblocks), dim3(256), 0, c.stream, n, x, c.y,
                           cp, c.Ap, c.Ai, c.diag, c.jtau, c.jmult, c.jmirror, c.jx, words, c.jmask, c.jwstart,
                           c.b, h, c.Araw, c.d_flags, d_total, (long long)c.cap);
    SD_HIP(hipGetLastError());
    c.araw_h = h;
    c.aux_valid = false;
    if (c.pc > 0) {
        hipLaunchKernelGGL(k_jac_pd, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, c.pc, x, c.y, c.b, c.px,
                           c.py, c.p.A, c.p.KASQR, cp, c.jpd);
        SD_HIP(hipGetLastError());
    }
    c.pattern_valid = true;
    c.pattern_slot = slot;
    c.jrc2 = cp.rc2;
    return SDDDST_OK;
}

static long long *jac_total(Context &c) { return reinterpret_cast<long long *>(c.d_flags + 2); }

// tau_x, the taper and the mirror positions of the current pattern (for a
// reuse at another step size and for the zero-value compaction): the full
// values pass again over the positions in slot (those the pattern was built
// from), which also rewrites the unsigned values for araw_h.
static int ensure_aux(Context &c, int slot)
{
    if (c.aux_valid) return SDDDST_OK;
    const int n = c.n;
    const int wblocks = (n + 3) / 4;
    const int words = (n + 63) / 64;
    const CutParams cp = cut_params(c);
    const double *x = c.x[slot];
    if (c.p.field_mode == SDDDST_FIELD_FAST)
        hipLaunchKernelGGL((k_jac_values<SDDDST_FIELD_FAST, true>), dim3(wblocks), dim3(256), 0, c.stream, n, x, c.y,
                           cp, c.Ap, c.Ai, c.diag, c.jtau, c.jmult, c.jmirror, c.jx, words, c.jmask, c.jwstart,
                           c.b, c.araw_h, c.Araw, c.d_flags, jac_total(c), (long long)c.cap);
    else
        hipLaunchKernelGGL((k_jac_values<SDDDST_FIELD_EXACT, true>), dim3(wblocks), dim3(256), 0, c.stream, n, x, c.y,
                           cp, c.Ap, c.Ai, c.diag, c.jtau, c.jmult, c.jmirror, c.jx, words, c.jmask, c.jwstart,
                           c.b, c.araw_h, c.Araw, c.d_flags, jac_total(c), (long long)c.cap);
    SD_HIP(hipGetLastError());
    c.aux_valid = true;
    return SDDDST_OK;
}

// reuse_pattern: the positions are those of the last build (stage II of a
// step); only the h-dependent passes run.  With h half the step size of the
// last unscaled values (stage II after stage I) those values and their
// column sums are reused times 0.5, which is exact (h enters every unscaled
// value and every partial sum as a factor, and halving commutes with
// rounding).
int launch_jacobian(Context &c, int slot, double h, bool reuse_pattern)
{
    const int n = c.n;
    const int wblocks = (n + 3) / 4;
    int rc;
    if ((c.timing & kTimeJacobian) && (rc = timer_begin(c, c.t_jac))) return rc;
    if (!reuse_pattern || !c.pattern_valid) {
        rc = build_pattern(c, slot, h);
        if (rc) return rc;
        c.raw_valid = false;
    }
    c.partner_valid = false;
    double factor = 1.0;
    if (reuse_pattern && c.raw_valid && c.jac_symmetric && h == 0.5 * c.raw_h) {
        factor = 0.5;
        hipLaunchKernelGGL(k_jac_dvec, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, c.sraw, factor, c.dvec,
                           c.colinfo);
        SD_HIP(hipGetLastError());
    } else {
        if (c.araw_h != h) {
            // the pattern is reused with another step size: the unsigned
            // values again, from tau_x and the taper (with the zero flag)
            if ((rc = ensure_aux(c, slot))) return rc;
            hipLaunchKernelGGL(k_flags_reset, dim3(1), dim3(64), 0, c.stream, c.d_flags);
            hipLaunchKernelGGL(k_jac_raw, dim3(wblocks), dim3(256), 0, c.stream, n, c.b, h, c.Ap, c.Ai, c.diag,
                               c.jtau, c.jmult, c.jmirror, c.pc > 0 ? c.jpd : nullptr, c.Araw, c.d_flags);
            SD_HIP(hipGetLastError());
            c.araw_h = h;
        }
        hipLaunchKernelGGL(k_jac_rawsum, dim3(wblocks), dim3(256), 0, c.stream, n, h, c.Ap, c.diag,
                           c.pc > 0 ? c.jpd : nullptr, c.Araw, c.sraw, c.dvec, c.colinfo, jac_total(c),
                           (long long)c.cap, c.d_flags + 4);
        SD_HIP(hipGetLastError());
        // d_flags[0..1] as one word ([0]: the zero-value flag), then the
        // entry count of the last build, then [4]: the strongly shifted
        // diagonals
        long long words2[3] = {0, 0, 0};
        if ((rc = host_wait(c, c.d_flags, 3, words2))) return rc;
        const long long total = words2[1];
        {
            int w4[2];
            std::memcpy(w4, &words2[2], sizeof w4);
            c.diag_strong = w4[0];
            c.diag_strong_h = h;
        }
        SD_CHECK(total < (1LL << 31) - 1, SDDDST_ERR_ARG, "nnz %lld exceeds the int32 index range", total);
        if (total > c.cap) {
            // the pattern outgrew the storage: grow it and assemble again
            // (every kernel after the count skipped its work)
            if ((rc = ensure_capacity(c, total))) return rc;
            c.pattern_valid = false;
            return launch_jacobian(c, slot, h, false);
        }
        c.nnz = total;
        int flags2[2];
        std::memcpy(flags2, &words2[0], sizeof flags2);
        const int flag = flags2[0];
        c.jac_symmetric = flag == 0;
        c.raw_valid = flag == 0;
        c.raw_h = h;
        if (flag) {
            // an exactly zero value: its upper copy goes (simulation.cpp:263-267)
            if ((rc = ensure_aux(c, slot))) return rc;
            hipLaunchKernelGGL(k_jac_raw, dim3(wblocks), dim3(256), 0, c.stream, n, c.b, h, c.Ap, c.Ai, c.diag,
                               c.jtau, c.jmult, c.jmirror, c.pc > 0 ? c.jpd : nullptr, c.Ax, c.d_flags);
            SD_HIP(hipGetLastError());
            rc = compact_zero_upper(c);
            if (rc) return rc;
            c.pattern_valid = false;   // positions of the kept entries moved
            hipLaunchKernelGGL(k_jac_diag, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, c.Ap, c.diag, c.Ax,
                               c.dvec);
            hipLaunchKernelGGL(k_jac_scale, dim3(wblocks), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag, c.dvec,
                               c.Ax, c.AxRow);
            SD_HIP(hipGetLastError());
            c.ax_valid = true;
            if ((c.timing & kTimeJacobian) && (rc = timer_end(c, c.t_jac))) return rc;
            c.jac_valid = true;
            c.jac_h = h;
            return SDDDST_OK;
        }
    }
    // the 16-bit row indices, once per pattern
    unsigned short *i16 = (c.Ai16 && !c.ai16_valid) ? c.Ai16 : nullptr;
    // the row-major values only: the column-major ones follow on request
    // (ensure_ax); the step's solvers read the row-major ones
    hipLaunchKernelGGL(k_jac_scale_partner, dim3(wblocks), dim3(256), 0, c.stream, n, c.Ap, c.Ai, c.diag,
                       c.colinfo, c.Araw, factor, (double *)nullptr, c.AxRow, c.partner, c.strength, i16);
    SD_HIP(hipGetLastError());
    c.ax_valid = false;
    c.ax_factor = factor;
    if (c.Ai16) c.ai16_valid = true;
    c.partner_valid = true;
    if ((c.timing & kTimeJacobian) && (rc = timer_end(c, c.t_jac))) return rc;
    c.jac_valid = true;
    c.jac_h = h;
    return SDDDST_OK;
}

int ensure_ax(Context &c)
{
    if (c.ax_valid) return SDDDST_OK;
    SD_CHECK(c.jac_valid, SDDDST_ERR_STATE, "no Jacobian assembled");
    hipLaunchKernelGGL(k_jac_ax, dim3((c.n + 3) / 4), dim3(256), 0, c.stream, c.n, c.Ap, c.Ai, c.diag, c.colinfo,
                       c.Araw, c.ax_factor, c.Ax);
    SD_HIP(hipGetLastError());
    c.ax_valid = true;
    return SDDDST_OK;
}

}  // namespace sdamd
