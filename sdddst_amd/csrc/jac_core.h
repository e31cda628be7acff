// Per-column bodies of the cutoff-limited Jacobian assembly
// (Simulation::calculateJacobian, simulation.cpp:236-384), shared by the
// single-system kernels (jacobian.hip) and the batched ensemble
// (ensemble.hip).  Every function works on one column j of one system whose
// arrays start at the given pointers; the wave-level ones expect the 64
// lanes of a wavefront to call them together.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sdddst_amd.h"
#include "field.h"

namespace sdamd {
namespace jac {

struct CutParams {
    double rc, rc2, thr, prefilter, inv_rc2;
};

// Predicate constants for the cutoff rc (rc2 = rc*rc, inv_rc2 = 1/rc2 as
// SimulationData::updateCutOff stores them, simulation_data.cpp:208-218).
inline CutParams make_cut_params(double rc, double rc2, double inv_rc2)
{
    CutParams cp;
    cp.rc = rc;
    cp.rc2 = rc2;
    cp.thr = 36.8 * rc2;
    cp.inv_rc2 = inv_rc2;
    // r < (1 + sqrt(36.8)) rc is necessary for the predicate; the margin
    // covers every rounding of the exact test.
    const double rmax = (1.0 + 6.06630035524124) * rc;
    cp.prefilter = rmax * rmax * (1.0 + 1e-9) + 1e-300;
    if (!(cp.prefilter < 1e300)) cp.prefilter = __builtin_inf();
    return cp;
}

// simulation.cpp:335 for the pair (lo, hi), dx = x_hi - x_lo.
__device__ __forceinline__ bool pair_pred(double dx, double dy, const CutParams &cp, double &r2_out)
{
    const double r2 = dx * dx + dy * dy;
    r2_out = r2;
    if (!(r2 < cp.prefilter)) return false;
    const double s = sdlibm::sqrt(r2) - cp.rc;
    return s * s < cp.thr;
}

// Row `row` inside [lo, hi) of a column (rows ascending).
__device__ __forceinline__ int find_row(const int *__restrict__ Ai, int lo, int hi, int row)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = Ai[mid];
        if (v < row) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Membership bitmask of column j (mj[words]) and its entry count; the
// predicate runs once per ordered pair with a cheap r^2 pre-filter.  With
// cmj, also the bitmask of the off-diagonal rows whose pair lies within the
// cutoff (r^2 <= rc^2: the untapered entries) and its count into *ccnt_j.
__device__ __forceinline__ void count_column(int j, int n, int words, const double *__restrict__ x,
                                             const double *__restrict__ y, const CutParams &cp,
                                             int *__restrict__ cnt_j, unsigned long long *__restrict__ mj,
                                             unsigned long long *__restrict__ cmj = nullptr,
                                             int *__restrict__ ccnt_j = nullptr)
{
    const int lane = (int)(threadIdx.x & 63);
    const double xj = x[j], yj = y[j];
    int c = 0, cc = 0;
    for (int w = 0; w < words; ++w) {
        const int i = w * 64 + lane;
        bool in = false, core = false;
        if (i < n) {
            if (i == j) {
                in = true;
            } else {
                const double xi = x[i], yi = y[i];
                // dx = x_hi - x_lo for the pair {i, j}
                const double dx = sdfield::normalize(i > j ? xi - xj : xj - xi);
                const double dy = sdfield::normalize(i > j ? yi - yj : yj - yi);
                double r2;
                in = pair_pred(dx, dy, cp, r2);
                core = !(r2 > cp.rc2);
            }
        }
        const unsigned long long m = __ballot(in);
        if (lane == 0) mj[w] = m;
        c += __popcll(m);
        if (cmj) {
            const unsigned long long cm = __ballot(core);
            if (lane == 0) cmj[w] = cm;
            cc += __popcll(cm);
        }
    }
    if (lane == 0) *cnt_j = c;
    if (cmj && lane == 0) *ccnt_j = cc;
}

// Rows of column j in ascending order from its bitmask, starting at base;
// records the diagonal position.
// wstart_j (nullable) receives the position in Ai of the first row of each
// mask word of the column, so that the position of row r is
// wstart_j[r / 64] + popcount(word bits below r) (see row_position).
__device__ __forceinline__ void fill_column(int j, int words, const unsigned long long *__restrict__ mj, int base,
                                            int *__restrict__ Ai, int *__restrict__ diag_j,
                                            int *__restrict__ wstart_j = nullptr)
{
    const int lane = (int)(threadIdx.x & 63);
    for (int w0 = 0; w0 < words; w0 += 64) {
        const int w = w0 + lane;
        const unsigned long long m = w < words ? mj[w] : 0ull;
        const int pc = __popcll(m);
        // exclusive prefix of pc across the wave
        int incl = pc;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        int pos = base + incl - pc;
        if (wstart_j && w < words) wstart_j[w] = pos;
        unsigned long long bits = m;
        while (bits) {
            const int bit = __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            const int i = w * 64 + bit;
            Ai[pos] = i;
            if (i == j) *diag_j = pos;
            ++pos;
        }
        base += __shfl(incl, 63, 64);
    }
}

// Position of row r in column i from the column's mask words and word
// starts (the same answer as find_row over the column's rows).
__device__ __forceinline__ int row_position(int words, const unsigned long long *__restrict__ mask,
                                            const int *__restrict__ wstart, int i, int r)
{
    const size_t w = (size_t)i * words + (r >> 6);
    const unsigned long long below = (1ull << (r & 63)) - 1ull;
    return wstart[w] + __popcll(mask[w] & below);
}

// tau_x and the taper of every strictly-lower entry of column j (row i > j)
// and the position of its mirror (row j in column i); with xsnap, x_j is kept
// there (the positions the pattern was built from, from which a later pass
// tells the entries within the cutoff).  Mode FAST takes the production derivative
// (xy_diff_x_fast), the other modes the reference's expression.  With raw,
// also the unsigned value for step size h at the entry and at its mirror.
// Without Aux, tau_x, the taper and the mirror positions are not kept (a
// later call that needs them runs the pass again with Aux).
template <int Mode, bool Aux = true>
__device__ __forceinline__ void values_column(int j, const double *__restrict__ x, const double *__restrict__ y,
                                              const CutParams &cp, const int *__restrict__ Ap,
                                              const int *__restrict__ Ai, const int *__restrict__ diag,
                                              double *__restrict__ tau, double *__restrict__ mult,
                                              int *__restrict__ mirror, double *__restrict__ xsnap,
                                              int words = 0, const unsigned long long *__restrict__ mask = nullptr,
                                              const int *__restrict__ wstart = nullptr,
                                              const double *__restrict__ b = nullptr, double h = 0.0,
                                              double *__restrict__ raw = nullptr, int *__restrict__ flags = nullptr)
{
    const int lane = (int)(threadIdx.x & 63);
    const int d = diag[j];
    const int end = Ap[j + 1];
    const double xj = x[j], yj = y[j];
    const double bj = raw ? b[j] : 0.0;
    if (xsnap && lane == 0) xsnap[j] = xj;
    for (int k = d + 1 + lane; k < end; k += 64) {
        const int i = Ai[k];
        // mirror: row j in column i (two loads with the word starts, else a
        // binary search of the column)
        const int mk = wstart ? row_position(words, mask, wstart, i, j) : find_row(Ai, Ap[i], diag[i], j);
        const double dx = sdfield::normalize(x[i] - xj);
        const double dy = sdfield::normalize(y[i] - yj);
        const double r2 = dx * dx + dy * dy;
        double m = 1;
        if (r2 > cp.rc2) {
            const double s = sdlibm::sqrt(r2) - cp.rc;
            m = sdlibm::exp(-(s * s) * cp.inv_rc2);
        }
        double t;
        if constexpr (Mode == SDDDST_FIELD_FAST) t = sdfield::xy_diff_x_fast(dx, dy);
        else t = sdfield::xy_diff_x(dx, dy);
        if constexpr (Aux) {
            tau[k] = t;
            mult[k] = m;
            mirror[k] = mk;
        }
        if (raw) {
            // the unsigned value h * b_i * b_j * tau * taper in the operation
            // order of simulation.cpp:337-343; the upper copy (row j of column
            // i, simulation.cpp:260-268) is the same product.  An exact zero
            // is flagged: its upper copy must not exist.
            const double v = h * b[i] * bj * t * m;
            raw[k] = v;
            raw[mk] = v;
            if (v == 0.0) atomicOr(flags, 1);
        }
    }
}

// simulation.cpp:325, 343: raw value h*b_i*b_j*tau_x*taper at the entry and
// at its mirror (simulation.cpp:260-268); the diagonal holds -pd*h.  Exact
// zeros are flagged: their mirrors must not exist.
__device__ __forceinline__ void raw_column(int j, const double *__restrict__ b, double h, const int *__restrict__ Ap,
                                           const int *__restrict__ Ai, const int *__restrict__ diag,
                                           const double *__restrict__ tau, const double *__restrict__ mult,
                                           const int *__restrict__ mirror, double pd, double *__restrict__ Ax,
                                           int *__restrict__ flags)
{
    const int lane = (int)(threadIdx.x & 63);
    const int d = diag[j];
    const int end = Ap[j + 1];
    const double bj = b[j];
    if (lane == 0) Ax[d] = -pd * h;
    for (int k = d + 1 + lane; k < end; k += 64) {
        const double v = h * b[Ai[k]] * bj * tau[k] * mult[k];
        Ax[k] = v;
        Ax[mirror[k]] = v;
        if (v == 0.0) atomicOr(flags, 1);
    }
}

// simulation.cpp:349-374 (one lane per column): the column sum in storage
// order, the diagonal s and d_j.  Eight loads are issued before their adds
// so the memory latency overlaps; the adds keep the sequential order.
__device__ __forceinline__ void diag_column(int j, const int *__restrict__ Ap, const int *__restrict__ diag,
                                            double *__restrict__ Ax, double *__restrict__ dvec)
{
    double subSum = 0;
    const int e = Ap[j + 1];
    int k = Ap[j];
    for (; k + 8 <= e; k += 8) {
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = Ax[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) subSum += a[u];
    }
    for (; k < e; ++k) subSum += Ax[k];
    subSum *= -1.;
    Ax[diag[j]] = subSum;
    double dv;
    if (subSum > 0) {
        subSum = 1. / subSum;
        subSum += 1.;
        subSum *= subSum;
        dv = 1. / subSum;
    } else {
        dv = 0.;
    }
    dvec[j] = dv;
}

// simulation.cpp:376-383: A = row-scaled values (+1 on the diagonal) and the
// same matrix row-major (AxRow: row j of A on the index arrays of column j).
__device__ __forceinline__ void scale_column(int j, const int *__restrict__ Ap, const int *__restrict__ Ai,
                                             const int *__restrict__ diag, const double *__restrict__ dvec,
                                             double *__restrict__ Ax, double *__restrict__ AxRow)
{
    const int lane = (int)(threadIdx.x & 63);
    const double Dj = (1.0 + dvec[j]) * 0.5;
    const int dpos = diag[j];
    for (int k = Ap[j] + lane; k < Ap[j + 1]; k += 64) {
        const double raw = Ax[k];
        double a = raw * ((1.0 + dvec[Ai[k]]) * 0.5);
        double r = raw * Dj;
        if (k == dpos) {
            a += 1.0;
            r += 1.0;
        }
        Ax[k] = a;
        if (AxRow) AxRow[k] = r;
    }
}

}  // namespace jac
}  // namespace sdamd
