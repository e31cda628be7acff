// Row bodies of the pair-block preconditioner, shared by the single-system
// FGMRES (solver_fgmres.hip) and the batched ensemble (ensemble.hip).  Row i
// of A lives at positions Ap[i]..Ap[i+1] of the symmetric pattern with
// A_ij = AxRow[k] and A_ji = Ax[k].
#pragma once

#include <hip/hip_runtime.h>

namespace sdamd {
namespace pairblk {

// Partner of row i: the j maximising |A_ij A_ji| / |A_ii A_jj| (ties to the
// smaller j); all 64 lanes of a wavefront call it together.
__device__ __forceinline__ void partner_row(int i, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                            const int *__restrict__ diag, const double *__restrict__ Ax,
                                            const double *__restrict__ AxRow, int *__restrict__ partner,
                                            double *__restrict__ strength)
{
    const int lane = (int)(threadIdx.x & 63);
    const double aii = __builtin_fabs(AxRow[diag[i]]);
    double best = -1.0;
    int bj = -1;
    for (int k = Ap[i] + lane; k < Ap[i + 1]; k += 64) {
        const int j = Ai[k];
        if (j == i) continue;
        const double ajj = __builtin_fabs(AxRow[diag[j]]);
        const double s = __builtin_fabs(AxRow[k] * Ax[k]) / (aii * ajj);
        if (s > best || (s == best && j < bj)) {
            best = s;
            bj = j;
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
        partner[i] = bj;
        strength[i] = best;
    }
}

// Mutual partners with coupling above the threshold form a 2x2 block; the
// row of the block inverse is stored as (binv_ii, binv_ij, j), else 1/A_ii.
// Without Ax (symmetric pattern), A_ji is read as row j's entry at column i
// (AxRow in column j): the same value, since both are the pair's unsigned
// value times row j's scale.
__device__ __forceinline__ int find_in_column(const int *__restrict__ Ap, const int *__restrict__ Ai, int col,
                                              int row)
{
    int lo = Ap[col], hi = Ap[col + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Ai[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void block_row(int i, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                          const int *__restrict__ diag, const double *__restrict__ Ax,
                                          const double *__restrict__ AxRow, const int *__restrict__ partner,
                                          const double *__restrict__ strength, double threshold,
                                          double *__restrict__ binv, int *__restrict__ bcol)
{
    const double aii = AxRow[diag[i]];
    const int j = partner[i];
    if (j >= 0 && partner[j] == i && strength[i] > threshold) {
        const int lo = find_in_column(Ap, Ai, i, j);
        const double aij = AxRow[lo], ajj = AxRow[diag[j]];
        const double aji = Ax ? Ax[lo] : AxRow[find_in_column(Ap, Ai, j, i)];
        const double det = aii * ajj - aij * aji;
        binv[2 * i] = ajj / det;
        binv[2 * i + 1] = -aij / det;
        bcol[i] = j;
    } else {
        binv[2 * i] = 1.0 / aii;
        binv[2 * i + 1] = 0.0;
        bcol[i] = -1;
    }
}

}  // namespace pairblk
}  // namespace sdamd
