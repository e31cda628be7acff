// K1 tile body: the symmetric pair interactions of one 64 x 64 tile of one
// system (Simulation::calculateSpeeds, simulation.cpp:86-131), shared by the
// single-system kernel (speeds.hip) and the batched ensemble (ensemble.hip).
//
// One wavefront per tile (bi <= bj).  Lane l holds target i = 64 bi + l; the
// source tile sits in LDS and lane l visits source m = (l + s) mod 64 at
// sweep s, so the 64 lanes read 64 different LDS words.  Each pair term
// tau(x_i - x_j) (lower index first) is added to the row accumulator of i
// and subtracted from the column accumulator, which is rotated one lane per
// sweep so that it arrives back at lane m's owner after 64 sweeps.  Row and
// column sums land in partial[bj][i] and partial[bi][j]: every target
// receives exactly one partial per tile column, summed later in tile order,
// so the result does not depend on scheduling.  minr2 takes the minimum r^2
// per target over the pairs with r^2 < r2_keep: the tolerance update
// (precision_handler.cpp:72-87) acts only when r^2 * 0.25 * 1e-2 is below
// prec^2, so with r2_keep a little above 400 prec^2 the kept minimum gives
// the same update as the minimum over all pairs, and these rare pairs are
// recorded with an atomic min instead of a rotated accumulator.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sdddst_amd.h"
#include "field.h"

namespace sdamd {
namespace pairtile {

constexpr int kTile = 64;
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

__device__ __forceinline__ double rotate_down1(double v, int lane)
{
    // lane l receives the value of lane (l + 1) mod 64
    return __shfl(v, (lane + 1) & 63, 64);
}

template <int Mode>
__device__ __forceinline__ double field_tau(double dx, double dy)
{
    if constexpr (Mode == SDDDST_FIELD_FAST) return sdfield::xy_fast(dx, dy);
    else return sdfield::xy(dx, dy);
}

// tile id -> (bi, bj), 0 <= bi <= bj < nb, enumerated column by column.
__device__ __forceinline__ void tile_coords(int t, int &bi, int &bj)
{
    int c = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)c * (c + 1) / 2 > t) --c;
    while ((long)(c + 1) * (c + 2) / 2 <= t) ++c;
    bj = c;
    bi = t - c * (c + 1) / 2;
}

// Called by one whole 64-lane workgroup.
template <int Mode>
__device__ __forceinline__ void tile(int bi, int bj, int n, const double *__restrict__ x,
                                     const double *__restrict__ y, const double *__restrict__ b,
                                     double *__restrict__ partial, unsigned long long *__restrict__ minr2,
                                     double r2_keep)
{
    __shared__ double sx[kTile], sy[kTile], sb[kTile];
    const int l = (int)threadIdx.x;
    const int i = bi * kTile + l;
    const int jl = bj * kTile + l;
    sx[l] = jl < n ? x[jl] : 0.0;
    sy[l] = jl < n ? y[jl] : 0.0;
    sb[l] = jl < n ? b[jl] : 0.0;
    const double xi = i < n ? x[i] : 0.0;
    const double yi = i < n ? y[i] : 0.0;
    const double bI = i < n ? b[i] : 0.0;
    __syncthreads();
    const bool diag = (bi == bj);
    double row = 0.0, col = 0.0;
    for (int s = 0; s < kTile; ++s) {
        const int m = (l + s) & 63;
        const int j = bj * kTile + m;
        const bool valid = (i < n) && (j < n) && (!diag || l < m);
        const double dx = sdfield::normalize(xi - sx[m]);
        const double dy = sdfield::normalize(yi - sy[m]);
        const double t = bI * sb[m] * field_tau<Mode>(dx, dy);
        const double r2 = dx * dx + dy * dy;
        if (valid) {
            row += t;
            col -= t;
            if (r2 < r2_keep) {
                const unsigned long long bits = (unsigned long long)__double_as_longlong(r2);
                atomicMin(&minr2[i], bits);
                atomicMin(&minr2[j], bits);
            }
        }
        col = rotate_down1(col, l);
    }
    if (diag) {
        if (i < n) partial[(size_t)bi * n + i] = row + col;
    } else {
        if (i < n) partial[(size_t)bj * n + i] = row;
        if (jl < n) partial[(size_t)bi * n + jl] = col;
    }
}

// Strict upper triangle: tile id -> (bi, bj), 0 <= bi < bj < nb, column by
// column.
__device__ __forceinline__ void strict_coords(int t, int &bi, int &bj)
{
    int c = (int)((1.0 + sqrt(8.0 * (double)t + 1.0)) * 0.5);
    while ((long)c * (c - 1) / 2 > t) --c;
    while ((long)(c + 1) * c / 2 <= t) ++c;
    bj = c;
    bi = (int)(t - (long)c * (c - 1) / 2);
}

// A diagonal tile in half the sweeps.  Lane l meets m = (l + s) mod 64 for
// s = 1..32: every unordered pair of the tile once (s = 32 pairs l with
// l + 32, kept for l < 32).  The pair term is evaluated lower index first,
// as in the off-diagonal tiles: when m < l the difference is taken as
// x_m - x_l and the term enters the row with a minus sign and the column
// with a plus sign.  Called by one whole 64-lane workgroup.
template <int Mode>
__device__ __forceinline__ void diag_tile(int bk, int n, const double *__restrict__ x, const double *__restrict__ y,
                                          const double *__restrict__ b, double *__restrict__ partial,
                                          unsigned long long *__restrict__ minr2, double r2_keep)
{
    __shared__ double sx[kTile], sy[kTile], sb[kTile];
    const int l = (int)threadIdx.x;
    const int i = bk * kTile + l;
    const double xi = i < n ? x[i] : 0.0;
    const double yi = i < n ? y[i] : 0.0;
    const double bI = i < n ? b[i] : 0.0;
    sx[l] = xi;
    sy[l] = yi;
    sb[l] = bI;
    __syncthreads();
    double row = 0.0, col = 0.0;
    for (int s = 1; s <= kTile / 2; ++s) {
        const int m = (l + s) & 63;
        const int j = bk * kTile + m;
        const bool lower = m > l;
        const bool valid = (i < n) && (j < n) && (s < kTile / 2 || l < kTile / 2);
        const double ex = lower ? xi - sx[m] : sx[m] - xi;
        const double ey = lower ? yi - sy[m] : sy[m] - yi;
        const double dx = sdfield::normalize(ex);
        const double dy = sdfield::normalize(ey);
        const double t = bI * sb[m] * field_tau<Mode>(dx, dy);
        const double r2 = dx * dx + dy * dy;
        if (valid) {
            const double tl = lower ? t : -t;
            row += tl;
            col -= tl;
            if (r2 < r2_keep) {
                const unsigned long long bits = (unsigned long long)__double_as_longlong(r2);
                atomicMin(&minr2[i], bits);
                atomicMin(&minr2[j], bits);
            }
        }
        col = rotate_down1(col, l);
    }
    // after sweeps 1..32 the column sum of source m sits in lane (m - 33) mod 64
    const double cs = __shfl(col, (l + 31) & 63, 64);
    if (i < n) partial[(size_t)bk * n + i] = row + cs;
}

// r^2 below which a pair can change a tolerance: 400 prec^2 with a margin
// for the roundings of r^2 * 0.25 * 1e-2.
__host__ __device__ inline double tolerance_r2(double minPrecSqr) { return minPrecSqr * 400.0 * (1.0 + 1e-6); }

}  // namespace pairtile
}  // namespace sdamd
