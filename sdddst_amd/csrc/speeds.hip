// K1: pair shear-stress speeds (Simulation::calculateSpeeds,
// simulation.cpp:86-131) and the tolerance update folded into it
// (PrecisionHandler::updateTolerance, precision_handler.cpp:72-87).
//
// Layout: x, y, b are SoA FP64 arrays of N.  The dislocations are ordered
// along a Hilbert curve over the unit cell (perm, a stable radix sort of the
// curve index of the wrapped (x, y), refreshed every 16 calls and after an
// upload): tile positions are positions in that order, so most tiles hold
// only pairs far apart.  The N(N-1)/2 unordered pairs are cut into 64x64 tiles
// (bi <= bj) of the upper triangle; one wavefront owns one tile.  Lane l
// holds target p = 64*bi + l in registers and walks the 64 sources of block
// bj in a rotated order (lane l meets source (l + s) mod 64 at step s), so
// every lane computes a distinct pair each step and each pair term
// t = b_i b_j tau(x_i - x_j, y_i - y_j) (i < j in the original numbering) is
// evaluated once:
//   - the row sum of the target stays in a register;
//   - the column sum of the source travels with it: the accumulator rotates
//     one lane per step (ds_bpermute), and after 64 steps lane l holds the
//     column sum of source 64*bj + l.
// Each tile writes its row sums into slot bj and its column sums into slot bi
// of a [nb][N] partial buffer at the sorted positions (a diagonal tile writes
// row + column into slot bi), so every (slot, position) is written exactly
// once; the final kernel adds the nb slots of a position in slot order and
// stores the speed of dislocation perm[position].
// The result is deterministic (no atomics on sums, an order that depends on
// the positions only) and independent of scheduling.
//
// The minimum r^2 per target (for the tolerance) is exact under any order
// and is combined with 64-bit atomic min on the bit pattern (r^2 >= 0).
//
// The ordered kernel (SDDDST_FIELD_ORDERED) runs one lane per target and
// sums in the reference's order: -t(i', i) for i' < i ascending, then
// +t(i, j) for j > i ascending.  It evaluates every pair twice and serves
// validation: with exact pair terms it reproduces the reference's speeds.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "context.h"
#include "far_expand.h"
#include "field.h"
#include "pair_core.h"
#include "step_core.h"

namespace sdamd {

namespace {

using pairtile::kInfBits;
using pairtile::kTile;

// Work units of equal cost: the nb (nb - 1) / 2 off-diagonal tiles, then
// the diagonal tiles two to a unit (each in half the sweeps), so that the
// units divide evenly over the resident waves.
// Tile positions follow perm (dislocations ordered along the curve).
template <int Mode>
__global__ __launch_bounds__(64) void k_pair_tiles(int n, int nb, int nfull, const double *__restrict__ x,
                                                   const double *__restrict__ y, const double *__restrict__ b,
                                                   const int *__restrict__ perm, double *__restrict__ partial,
                                                   unsigned long long *__restrict__ minr2, double r2_keep,
                                                   pairtile::FarData fd)
{
    const int u = (int)blockIdx.x;
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        // an x-separated tile is summed by its series (k_farx_gather): its
        // partial sums are zero here
        if (fd.expand && u < nfull) {
            int bi, bj;
            pairtile::strict_coords(u, bi, bj);
            if (farx::tile_code(fd.ext + (size_t)bi * pairtile::kExtWords, fd.ext + (size_t)bj * pairtile::kExtWords)) {
                const int l = (int)threadIdx.x;
                const int p = bi * kTile + l, q = bj * kTile + l;
                if (p < n) partial[(size_t)bj * n + p] = 0.0;
                if (q < n) partial[(size_t)bi * n + q] = 0.0;
                return;
            }
        }
    }
    __shared__ pairtile::TileLds L;
    // The fast field's table lookups (cos and exp, per lane at data-dependent
    // indices) read an LDS copy: 32-bit LDS addressing instead of 64-bit
    // global address arithmetic per lookup.  The copy is written here and
    // ordered before its first read by the tile's own barrier.
    const double *sct = sdlibm_sincos_tab;
    const uint64_t *e2t = sdlibm_exp2_tab;
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        constexpr int nsc = (int)(sizeof(sdlibm_sincos_tab) / sizeof(double));
        constexpr int ne2 = (int)(sizeof(sdlibm_exp2_tab) / sizeof(uint64_t));
        __shared__ double lsc[nsc];
        __shared__ uint64_t le2[ne2];
        for (int k = (int)threadIdx.x; k < nsc; k += kTile) lsc[k] = sdlibm_sincos_tab[k];
        for (int k = (int)threadIdx.x; k < ne2; k += kTile) le2[k] = sdlibm_exp2_tab[k];
        sct = lsc;
        e2t = le2;
    }
    if (u < nfull) {
        int bi, bj;
        pairtile::strict_coords(u, bi, bj);
        if ((bj + 1) * kTile <= n)
            pairtile::tile<Mode, true, true>(L, bi, bj, n, x, y, b, perm, partial, minr2, r2_keep, sct, e2t, fd);
        else pairtile::tile<Mode, true, false>(L, bi, bj, n, x, y, b, perm, partial, minr2, r2_keep, sct, e2t, fd);
        return;
    }
    const int bk = 2 * (u - nfull);
    pairtile::diag_tile<Mode, true>(L, bk, n, x, y, b, perm, partial, minr2, r2_keep, sct, e2t, fd);
    if (bk + 1 < nb) {
        __syncthreads();
        pairtile::diag_tile<Mode, true>(L, bk + 1, n, x, y, b, perm, partial, minr2, r2_keep, sct, e2t, fd);
    }
}

// Far-pair factors of the fast field: cos and sin of 2 pi y per dislocation
// (y wrapped first; y changes only by upload), and the wrapped x with
// e^{+-2 pi x} (per speeds call).
__global__ void k_ytrig(int n, const double *__restrict__ y, double *__restrict__ yc, double *__restrict__ ys,
                        double *__restrict__ yw)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double cv, sv;
    sdfield::ytrig(y[i], cv, sv);
    yc[i] = cv;
    ys[i] = sv;
    yw[i] = y[i] - __builtin_rint(y[i]);
}

// Sort keys of the pair-tile order: the index along a Hilbert curve of the
// cell that holds the wrapped (x, y) in a 2^kOrderBits x 2^kOrderBits grid
// of the unit cell (the curve's published iterative construction: per bit
// level the quadrant's rank, then the quadrant's rotation or reflection),
// and the index (a stable radix sort keeps equal cells in index order, so
// the order is a function of the positions alone).  Consecutive runs of 64
// along the curve are compact in both x and y, so most tile pairs are
// all-far.  The context's speeds use the strip order below (more tiles
// x-separated); the allocation-free entry point keeps the curve.
constexpr int kOrderBits = 10;
// strips = kStripFactor sqrt(blocks) (64 at N = 16384; measured on the
// headline workload: pair kernel 0.282 ms along the curve, 0.231 with 32
// strips, 0.224 with 64 or 96, 0.227 with 128); 0: the curve
constexpr double kStripFactor = 4.0;
__global__ void k_curvekey(int n, const double *__restrict__ x, const double *__restrict__ y,
                           int *__restrict__ key, int *__restrict__ idx)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    constexpr int g = 1 << kOrderBits;
    const double wx = x[i] - floor(x[i] + 0.5), wy = y[i] - floor(y[i] + 0.5);
    int cx = (int)((wx + 0.5) * (double)g), cy = (int)((wy + 0.5) * (double)g);
    cx = cx < 0 ? 0 : (cx >= g ? g - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= g ? g - 1 : cy);
    int d = 0;
    for (int s = g >> 1; s > 0; s >>= 1) {
        const int rx = (cx & s) ? 1 : 0, ry = (cy & s) ? 1 : 0;
        d += s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                cx = g - 1 - cx;
                cy = g - 1 - cy;
            }
            const int t = cx;
            cx = cy;
            cy = t;
        }
    }
    key[i] = d;
    idx[i] = i;
}

// Strip order: the wrapped x cut into `strips` equal strips, taken left to
// right, each walked in y up and down alternately (y in 2^kOrderBits
// levels); equal keys keep index order under the stable sort.  Runs of 64
// are then narrow in x, so more tiles are x-separated (summed by the series)
// than along the curve.
__global__ void k_stripkey(int n, int strips, const double *__restrict__ x, const double *__restrict__ y,
                           int *__restrict__ key, int *__restrict__ idx)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    constexpr int g = 1 << kOrderBits;
    const double wx = x[i] - floor(x[i] + 0.5), wy = y[i] - floor(y[i] + 0.5);
    int sx = (int)((wx + 0.5) * (double)strips), cy = (int)((wy + 0.5) * (double)g);
    sx = sx < 0 ? 0 : (sx >= strips ? strips - 1 : sx);
    cy = cy < 0 ? 0 : (cy >= g ? g - 1 : cy);
    key[i] = sx * g + ((sx & 1) ? g - 1 - cy : cy);
    idx[i] = i;
}

// simulation.cpp:111-128: point-defect force on dislocation i and its
// tolerance candidate (X2 + X2, simulation.cpp:119-126).
__device__ __forceinline__ double pd_pair(double xi, double yi, double bI, double pxj, double pyj, double A, double K,
                                          double &rSqr_out)
{
    const double dx = sdfield::normalize(xi - pxj);
    const double dy = sdfield::normalize(yi - pyj);
    const double pi = sdfield::kPi;
    const double xSqr = (1.0 - sdlibm::cos(2.0 * pi * dx)) * 0.5 / pi / pi;
    const double ySqr = (1.0 - sdlibm::cos(2.0 * pi * dy)) * 0.5 / pi / pi;
    const double rSqr = xSqr + ySqr;
    const double expXY = sdlibm::exp(-K * rSqr);
    const double Xdx = sdlibm::sin(2.0 * pi * dx) * 0.5 / pi;
    const double Xdy = sdlibm::sin(2.0 * pi * dy) * 0.5 / pi;
    rSqr_out = rSqr;
    return 2.0 * A * Xdx * Xdy * ((1. - expXY) / rSqr - K * expXY) / rSqr * bI;
}

// precision_handler.cpp:72-87 applied to the minimum candidate.
using stepf::tol_update;

__global__ void k_speeds_finish(int n, int nb, int pc, const int *__restrict__ perm,
                                const double *__restrict__ partial, const double *__restrict__ x,
                                const double *__restrict__ y, const double *__restrict__ b,
                                const double *__restrict__ px, const double *__restrict__ py, double A, double K,
                                double sigma, const unsigned long long *__restrict__ minr2,
                                double *__restrict__ tol, double minPrecSqr, int update_tol,
                                const double *__restrict__ vfar, double *__restrict__ v)
{
    // partial holds sorted positions; perm maps position q to dislocation i;
    // vfar (nullable) the share of the x-separated tiles at the same positions
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= n) return;
    const int i = perm[q];
    double acc = 0.0;
    int s = 0;
    for (; s + 8 <= nb; s += 8) {
        double p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = partial[(size_t)(s + u) * n + q];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += p[u];
    }
    for (; s < nb; ++s) acc += partial[(size_t)s * n + q];
    if (vfar) acc += vfar[q];
    double t = update_tol ? tol[i] : 0.0;
    if (update_tol) {
        const unsigned long long bits = minr2[i];
        if (bits != kInfBits) tol_update(t, __longlong_as_double((long long)bits), minPrecSqr);
    }
    if (pc > 0) {
        const double xi = x[i], yi = y[i], bI = b[i];
        for (int j = 0; j < pc; ++j) {
            double r2;
            acc -= pd_pair(xi, yi, bI, px[j], py[j], A, K, r2);
            if (update_tol) tol_update(t, r2, minPrecSqr);
        }
    }
    acc += b[i] * sigma;
    v[i] = acc;
    if (update_tol) tol[i] = t;
}

// One lane per target in the reference's summation order (validation).
__global__ void k_speeds_ordered(int n, int pc, const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ b, const double *__restrict__ px,
                                 const double *__restrict__ py, double A, double K, double sigma,
                                 double *__restrict__ tol, double minPrecSqr, int update_tol,
                                 double *__restrict__ v)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double xi = x[i], yi = y[i], bI = b[i];
    double acc = 0.0;
    double t = update_tol ? tol[i] : 0.0;
    for (int j = 0; j < i; ++j) {
        const double dx = sdfield::normalize(x[j] - xi);
        const double dy = sdfield::normalize(y[j] - yi);
        const double tmp = b[j] * bI * sdfield::xy(dx, dy);
        if (update_tol) tol_update(t, dx * dx + dy * dy, minPrecSqr);
        acc -= tmp;
    }
    for (int j = i + 1; j < n; ++j) {
        const double dx = sdfield::normalize(xi - x[j]);
        const double dy = sdfield::normalize(yi - y[j]);
        const double tmp = bI * b[j] * sdfield::xy(dx, dy);
        if (update_tol) tol_update(t, dx * dx + dy * dy, minPrecSqr);
        acc += tmp;
    }
    for (int j = 0; j < pc; ++j) {
        double r2;
        acc -= pd_pair(xi, yi, bI, px[j], py[j], A, K, r2);
        if (update_tol) tol_update(t, r2, minPrecSqr);
    }
    acc += bI * sigma;
    v[i] = acc;
    if (update_tol) tol[i] = t;
}

// Start of a speeds call: the tolerance candidates cleared, and with w (the
// fast field's far pairs) the wrapped x and e^{+-2 pi x} of each dislocation.
__global__ void k_speeds_prep(int n, unsigned long long *__restrict__ minr2, const double *__restrict__ x,
                              double *__restrict__ w, double *__restrict__ ep, double *__restrict__ em)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    minr2[i] = kInfBits;
    if (w) {
        double wv, pv, mv;
        sdfield::xexp(x[i], wv, pv, mv);
        w[i] = wv;
        ep[i] = pv;
        em[i] = mv;
    }
}

// The same for the fast field's far-pair path, in tile order: one
// wavefront per block of tile positions resets and forms its members'
// values (wrapped x, e^{+-2 pi w}) and the block's extents for the tile
// classification (pairtile::block_extents).
__global__ __launch_bounds__(64) void k_speeds_prep_blocks(int n, const int *__restrict__ perm,
                                                           unsigned long long *__restrict__ minr2,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ yw, double *__restrict__ w,
                                                           double *__restrict__ ep, double *__restrict__ em,
                                                           double *__restrict__ ext)
{
    const int p = (int)(blockIdx.x * 64 + threadIdx.x);
    double wv = __builtin_nan(""), yv = __builtin_nan("");
    if (p < n) {
        const int i = perm[p];
        minr2[i] = kInfBits;
        double pv, mv;
        sdfield::xexp(x[i], wv, pv, mv);
        w[i] = wv;
        ep[i] = pv;
        em[i] = mv;
        yv = yw[i];
    }
    double e[pairtile::kExtWords];
    pairtile::block_extents(wv, yv, e);
    if (threadIdx.x == 0) {
        double *o = ext + (size_t)blockIdx.x * pairtile::kExtWords;
#pragma unroll
        for (int k = 0; k < pairtile::kExtWords; ++k) o[k] = e[k];
    }
}

// Source moments of the x-separated tiles (far_expand.h), one workgroup of
// eight wavefronts per block of tile positions: the members' factors are
// staged in LDS, lane t - 1 of every wave takes term t over an eighth of the
// members (their powers by one binary exponentiation, the eight members in
// lockstep), and the eighths are added in order.
constexpr int kMomWaves = 8;
__global__ __launch_bounds__(64 * kMomWaves) void k_farx_moments(int n, const int *__restrict__ perm,
                                                                 const double *__restrict__ w,
                                                                 const double *__restrict__ ca,
                                                                 const double *__restrict__ sa,
                                                                 const double *__restrict__ b,
                                                                 const double *__restrict__ ext,
                                                                 double *__restrict__ mom, double *__restrict__ mom1)
{
    using namespace farx;
    static_assert(kTerms <= 64, "one lane per term");
    constexpr int W = kMomWaves, kPer = 64 / W, kLock = 8;
    static_assert(kPer % kLock == 0, "members in lockstep groups");
    __shared__ double sxi[64], sb[64], spr[64], spi[64], sr[64];
    __shared__ double part[W][64][kWords - 2];
    __shared__ double part1[W][kWords1];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int J = (int)blockIdx.x;
    const double X = centre(ext + (size_t)J * pairtile::kExtWords);
    if (wave == 0) {
        const int p = J * 64 + lane;
        double xi = 0.0, bb = 0.0, c = 1.0, s = 0.0;
        if (p < n) {
            const int j = perm[p];
            xi = w[j] - X;
            bb = b[j];
            c = ca[j];
            s = sa[j];
        }
        // phi+ = e^{2 pi xi} e^{-2 pi i y}; phi- = phi+ e^{-4 pi xi}
        const double ep = exp(sdfield::kTwoPi * xi), em = exp(-sdfield::kTwoPi * xi);
        sxi[lane] = xi;
        sb[lane] = bb;
        spr[lane] = c * ep;
        spi[lane] = -s * ep;
        sr[lane] = em * em;
    }
    __syncthreads();
    const int t = lane + 1;
    double v[kWords - 2] = {};
    double v1[kWords1] = {};
    if (t <= kTerms) {
        for (int m0 = wave * kPer; m0 < wave * kPer + kPer; m0 += kLock) {
            // phi+^t and e^{-4 pi t xi} of kLock members by one exponentiation
            Cplx a[kLock], pp[kLock];
            double r[kLock], rt[kLock];
#pragma unroll
            for (int u = 0; u < kLock; ++u) {
                a[u] = Cplx{spr[m0 + u], spi[m0 + u]};
                r[u] = sr[m0 + u];
                pp[u] = Cplx{1.0, 0.0};
                rt[u] = 1.0;
            }
            for (int k = t; k > 0; k >>= 1) {
                if (k & 1) {
#pragma unroll
                    for (int u = 0; u < kLock; ++u) {
                        pp[u] = cmul(pp[u], a[u]);
                        rt[u] *= r[u];
                    }
                }
#pragma unroll
                for (int u = 0; u < kLock; ++u) {
                    a[u] = cmul(a[u], a[u]);
                    r[u] *= r[u];
                }
            }
#pragma unroll
            for (int u = 0; u < kLock; ++u) {
                const Cplx pm = cscale(pp[u], rt[u]);
                const double xi = sxi[m0 + u], bb = sb[m0 + u], bx = bb * xi;
                v[0] += bb * pp[u].re;
                v[1] += bb * pp[u].im;
                v[2] += bx * pp[u].re;
                v[3] += bx * pp[u].im;
                v[4] += bb * pm.re;
                v[5] += bb * pm.im;
                v[6] += bx * pm.re;
                v[7] += bx * pm.im;
                if (t == 1) {
                    const double bxx = bx * xi;
                    v1[0] += bxx * pp[u].re;
                    v1[1] += bxx * pp[u].im;
                    v1[2] += bxx * pm.re;
                    v1[3] += bxx * pm.im;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kWords - 2; ++k) part[wave][lane][k] = v[k];
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < kWords1; ++k) part1[wave][k] = v1[k];
    __syncthreads();
    if (wave == 0 && t <= kTerms) {
        double *o = mom + ((size_t)J * kTerms + (t - 1)) * kWords;
        o[0] = exp(sdfield::kTwoPi * t * X);
        o[1] = exp(-sdfield::kTwoPi * t * X);
#pragma unroll
        for (int k = 0; k < kWords - 2; ++k) {
            double sum = part[0][lane][k];
            for (int q = 1; q < W; ++q) sum += part[q][lane][k];
            o[2 + k] = sum;
        }
        if (t == 1)
#pragma unroll
            for (int k = 0; k < kWords1; ++k) {
                double sum = part1[0][k];
                for (int q = 1; q < W; ++q) sum += part1[q][k];
                mom1[(size_t)J * kWords1 + k] = sum;
            }
    }
}

// Per target block I: the coefficients gathered over the source blocks of
// its x-separated tiles, then every member of the block sums its series:
// vfar[p] = b_i 4 pi^2 sum_n n Re(...), the tiles' share of the member's
// speed (the pair kernel skips them).  The blocks' x-extents are staged in
// LDS and the qualifying source blocks listed in ascending order first;
// wave v of eight takes list entries v, v + 8, ... with lane t - 1 on term
// t, and the eight partial coefficient sets are added in order.  For the series, wave v takes terms 8 v + 1 .. 8 v + 8
// of every member (lane m), the eight partial sums again added in order.
constexpr int kGatherWaves = 8;
constexpr size_t kGatherDynLds = 96 * 1024;   // with the static 53 KiB, within the CU's 160 KiB
__global__ __launch_bounds__(512) void k_farx_gather(int n, int nb, const int *__restrict__ perm,
                                                     const double *__restrict__ w, const double *__restrict__ ca,
                                                     const double *__restrict__ sa, const double *__restrict__ b,
                                                     const double *__restrict__ ext, const double *__restrict__ mom,
                                                     const double *__restrict__ mom1, double *__restrict__ vfar)
{
    using namespace farx;
    static_assert(kTerms == 64, "one lane per term, eight terms per wave in the series");
    constexpr int W = kGatherWaves;
    extern __shared__ double xe[];                         // [nb][2] x-extents, then the list
    int *list = reinterpret_cast<int *>(xe + 2 * (size_t)nb);   // J * 8 + code, J ascending
    __shared__ double acc[W][64][12];
    __shared__ double part[W][64];
    __shared__ int wcount[W];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int I = (int)blockIdx.x;
    for (int k = (int)threadIdx.x; k < nb; k += 64 * W) {
        xe[2 * k] = ext[(size_t)k * pairtile::kExtWords + 4];
        xe[2 * k + 1] = ext[(size_t)k * pairtile::kExtWords + 5];
    }
    __syncthreads();
    const double XI = 0.5 * (xe[2 * I] + xe[2 * I + 1]);
    int len = 0;
    for (int base = 0; base < nb; base += 64 * W) {
        const int J = base + (int)threadIdx.x;
        int code = 0;
        if (J < nb && J != I) {
            const int lo = I < J ? I : J, hi = I < J ? J : I;
            code = tile_code_x(xe[2 * lo], xe[2 * lo + 1], xe[2 * hi], xe[2 * hi + 1]);
        }
        const unsigned long long m = __ballot(code != 0);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int off = len;
        for (int v = 0; v < wave; ++v) off += wcount[v];
        if (code) list[off + __popcll(m & ((1ull << lane) - 1ull))] = J * 8 + code;
        for (int v = 0; v < W; ++v) len += wcount[v];
        __syncthreads();
    }
    const int t = lane + 1;
    // psi- and psi+ families, powers 0..2 of xi_i
    Cplx Lm0{0.0, 0.0}, Lm1{0.0, 0.0}, Lm2{0.0, 0.0}, Lp0{0.0, 0.0}, Lp1{0.0, 0.0}, Lp2{0.0, 0.0};
    {
        Poly alpha, beta;
        term_polys(t, alpha, beta);
        const double Gp = exp(sdfield::kTwoPi * t), Gm = exp(-sdfield::kTwoPi * t);
        const double *mI = mom + ((size_t)I * kTerms + (t - 1)) * kWords;
        const double EXI = mI[0], EXIinv = mI[1];
        const bool withB = t <= kTermsB;
        for (int e = wave; e < len; e += W) {
            const int J = list[e] >> 3, code = list[e] & 7;
            const double *mJ = mom + ((size_t)J * kTerms + (t - 1)) * kWords;
            const double EXJ = mJ[0], EXJinv = mJ[1];
            const Cplx M0p{mJ[2], mJ[3]}, M1p{mJ[4], mJ[5]}, M0m{mJ[6], mJ[7]}, M1m{mJ[8], mJ[9]};
            Cplx M2p{0.0, 0.0}, M2m{0.0, 0.0};
            if (t == 1) {
                const double *m1 = mom1 + (size_t)J * kWords1;
                M2p = Cplx{m1[0], m1[1]};
                M2m = Cplx{m1[2], m1[3]};
            }
            int sigma;
            double o;
            code_side(code, sigma, o);
            if (I > J) {
                sigma = -sigma;
                o = -o;
            }
            const double D = XI - 0.5 * (xe[2 * J] + xe[2 * J + 1]) + o;
            // e^{-+2 pi t o}
            const double eo_m = o > 0 ? Gm : (o < 0 ? Gp : 1.0);
            const double eo_p = o > 0 ? Gp : (o < 0 ? Gm : 1.0);
            if (sigma > 0) {
                // e^{-2 pi t D}: alpha terms, psi- with the + moments
                gather(Lm0, Lm1, Lm2, EXIinv * EXJ * eo_m, D, alpha, 1.0, -1.0, M0p, M1p, M2p);
                // e^{-2 pi t (1 - D)}: beta terms, psi+ with the - moments
                if (withB) gather(Lp0, Lp1, Lp2, Gm * EXI * EXJinv * eo_p, D, beta, 1.0, -1.0, M0m, M1m, M2m);
            } else {
                // |dx| = D' - xi_i + xi_j, D' = -D, tau = -tau(|dx|)
                gather(Lp0, Lp1, Lp2, -(EXI * EXJinv * eo_p), -D, alpha, -1.0, 1.0, M0m, M1m, M2m);
                if (withB) gather(Lm0, Lm1, Lm2, -(Gm * EXIinv * EXJ * eo_m), -D, beta, -1.0, 1.0, M0p, M1p, M2p);
            }
        }
    }
    {
        double *a = acc[wave][lane];
        a[0] = Lm0.re; a[1] = Lm0.im; a[2] = Lm1.re; a[3] = Lm1.im; a[4] = Lm2.re; a[5] = Lm2.im;
        a[6] = Lp0.re; a[7] = Lp0.im; a[8] = Lp1.re; a[9] = Lp1.im; a[10] = Lp2.re; a[11] = Lp2.im;
    }
    __syncthreads();
    for (int k = (int)threadIdx.x; k < 64 * 12; k += 64 * W) {
        const int l = k / 12, f = k % 12;
        double sum = acc[0][l][f];
        for (int v = 1; v < W; ++v) sum += acc[v][l][f];
        acc[0][l][f] = sum;
    }
    __syncthreads();
    // the block's members: lane m, terms 8 wave + 1 .. 8 wave + 8
    const int p = I * 64 + lane;
    double s = 0.0, bI = 0.0;
    if (p < n) {
        const int i = perm[p];
        bI = b[i];
        const double xi = w[i] - XI;
        const Cplx cy{ca[i], sa[i]};
        const Cplx gm = cscale(cy, exp(-sdfield::kTwoPi * xi));   // psi-
        const Cplx gp = cscale(cy, exp(sdfield::kTwoPi * xi));    // psi+
        constexpr int kQ = kTerms / W;
        const int t0 = wave * kQ + 1;
        Cplx pm = cpow(gm, t0), pp = cpow(gp, t0);
        for (int tt = t0; tt < t0 + kQ; ++tt) {
            const double *a = acc[0][tt - 1];
            const Cplx Lm0{a[0], a[1]}, Lm1{a[2], a[3]}, Lm2{a[4], a[5]};
            const Cplx Lp0{a[6], a[7]}, Lp1{a[8], a[9]}, Lp2{a[10], a[11]};
            const Cplx cm = cadd(Lm0, cscale(cadd(Lm1, cscale(Lm2, xi)), xi));
            const Cplx cp = cadd(Lp0, cscale(cadd(Lp1, cscale(Lp2, xi)), xi));
            const double re = (pm.re * cm.re - pm.im * cm.im) + (pp.re * cp.re - pp.im * cp.im);
            s += (double)tt * re;
            pm = cmul(pm, gm);
            pp = cmul(pp, gp);
        }
    }
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && p < n) {
        double sum = part[0][lane];
        for (int v = 1; v < W; ++v) sum += part[v][lane];
        vfar[p] = bI * (4.0 * sdfield::kPiSq) * sum;
    }
}

}  // namespace

int pair_order_alloc(Context &c)
{
    const int n = c.n;
    for (int **p : {&c.perm, &c.xkey, &c.xkey_sorted, &c.xidx})
        SD_HIP(hipMalloc((void **)p, sizeof(int) * (size_t)(n > 0 ? n : 1)));
    size_t bytes = 0;
    SD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c.xkey, c.xkey_sorted, c.xidx, c.perm, n, 0,
                                              2 * kOrderBits, c.stream));
    c.sort_tmp_bytes = bytes > 0 ? bytes : 1;
    SD_HIP(hipMalloc(&c.sort_tmp, c.sort_tmp_bytes));
    for (double **p : {&c.ytc, &c.yts, &c.yw, &c.xw, &c.xep, &c.xem})
        SD_HIP(hipMalloc((void **)p, sizeof(double) * (size_t)(n > 0 ? n : 1)));
    SD_HIP(hipMalloc((void **)&c.bext, sizeof(double) * pairtile::kExtWords * (size_t)(c.nb > 0 ? c.nb : 1)));
    const size_t nbb = (size_t)(c.nb > 0 ? c.nb : 1);
    SD_HIP(hipMalloc((void **)&c.farx_mom, sizeof(double) * farx::kTerms * farx::kWords * nbb));
    SD_HIP(hipMalloc((void **)&c.farx_mom1, sizeof(double) * farx::kWords1 * nbb));
    SD_HIP(hipMalloc((void **)&c.vfar, sizeof(double) * (size_t)(n > 0 ? n : 1)));
    return SDDDST_OK;
}

int launch_ytrig(Context &c)
{
    if (c.n <= 0) return SDDDST_OK;
    hipLaunchKernelGGL(k_ytrig, dim3((c.n + 255) / 256), dim3(256), 0, c.stream, c.n, c.y, c.ytc, c.yts, c.yw);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

constexpr int kOrderReuse = 16;

// perm = the dislocations in strip order (index order within equal keys).
static int pair_order(Context &c, const double *x)
{
    const int n = c.n;
    if (kStripFactor > 0) {
        const int strips = std::min(1 << kOrderBits, std::max(1, (int)lround(kStripFactor * sqrt((double)c.nb))));
        hipLaunchKernelGGL(k_stripkey, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, strips, x, c.y, c.xkey,
                           c.xidx);
    } else {
        hipLaunchKernelGGL(k_curvekey, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, x, c.y, c.xkey, c.xidx);
    }
    SD_HIP(hipGetLastError());
    size_t bytes = c.sort_tmp_bytes;
    SD_HIP(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp, bytes, c.xkey, c.xkey_sorted, c.xidx, c.perm, n, 0,
                                              2 * kOrderBits, c.stream));
    return SDDDST_OK;
}

int launch_speeds(Context &c, int slot, int out_buf, double sigma, int update_tol)
{
    const int n = c.n;
    const double minPrecSqr = c.p.precision * c.p.precision;
    const int threads = 256;
    const int blocks = (n + threads - 1) / threads;
    int rc;
    if ((c.timing & kTimeSpeeds) && (rc = timer_begin(c, c.t_speeds))) return rc;
    if (c.p.field_mode == SDDDST_FIELD_ORDERED) {
        if ((c.timing & kTimePair) && (rc = timer_begin(c, c.t_pair))) return rc;
        hipLaunchKernelGGL(k_speeds_ordered, dim3(blocks), dim3(threads), 0, c.stream, n, c.pc, c.x[slot], c.y, c.b,
                           c.px, c.py, c.p.A, c.p.KASQR, sigma, c.tol, minPrecSqr, update_tol, c.v[out_buf]);
        SD_HIP(hipGetLastError());
        if ((((c.timing & kTimePair) && (rc = timer_end(c, c.t_pair))) || ((c.timing & kTimeSpeeds) && (rc = timer_end(c, c.t_speeds))))) return rc;
        return SDDDST_OK;
    }
    const int nb = c.nb;
    const long nfull = (long)nb * (nb - 1) / 2;
    const long units = nfull + (nb + 1) / 2;
    SD_CHECK(units < (1L << 31), SDDDST_ERR_ARG, "too many tiles for N=%d", n);
    // the order is refreshed every kOrderReuse calls: positions move far less
    // than a bin width between them, and any order gives the same terms
    if (c.order_age == 0 || c.order_age >= kOrderReuse) {
        if ((rc = pair_order(c, c.x[slot]))) return rc;
        c.order_age = 0;
    }
    ++c.order_age;
    // far-pair factors (fast field): every pair whose gap cosh(2 pi dx) -
    // cos(2 pi dy) is at least kFarMinGap takes the far form, which skips the
    // tolerance candidates, so they are used only when no such pair can be
    // one (r^2 > kFarMinR2 for all of them)
    pairtile::FarData fd;
    const double r2_keep = pairtile::tolerance_r2(minPrecSqr);
    const bool far = c.p.field_mode == SDDDST_FIELD_FAST && r2_keep < sdfield::kFarMinR2;
    if (far)
        hipLaunchKernelGGL(k_speeds_prep_blocks, dim3(nb), dim3(kTile), 0, c.stream, n, c.perm, c.minr2, c.x[slot],
                           c.yw, c.xw, c.xep, c.xem, c.bext);
    else
        hipLaunchKernelGGL(k_speeds_prep, dim3(blocks), dim3(threads), 0, c.stream, n, c.minr2, c.x[slot], nullptr,
                           c.xep, c.xem);
    SD_HIP(hipGetLastError());
    if (far) {
        fd.ca = c.ytc;
        fd.sa = c.yts;
        fd.w = c.xw;
        fd.ep = c.xep;
        fd.em = c.xem;
        fd.yw = c.yw;
        fd.ext = c.bext;
        // (the series' gather stages 20 bytes per block in LDS)
        fd.expand = nb >= 2 && (2 * sizeof(double) + sizeof(int)) * (size_t)nb <= kGatherDynLds ? 1 : 0;
    }
    if (fd.expand) {
        // the gather's LDS: 53 KiB of its own plus 20 bytes per block (above
        // the default limit for the larger systems)
        static std::once_flag lds_once;
        static hipError_t lds_rc = hipSuccess;
        std::call_once(lds_once, [] {
            lds_rc = hipFuncSetAttribute((const void *)k_farx_gather, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)kGatherDynLds);
        });
        SD_HIP(lds_rc);
        hipLaunchKernelGGL(k_farx_moments, dim3(nb), dim3(64 * kMomWaves), 0, c.stream, n, c.perm, c.xw, c.ytc,
                           c.yts, c.b, c.bext, c.farx_mom, c.farx_mom1);
        hipLaunchKernelGGL(k_farx_gather, dim3(nb), dim3(64 * kGatherWaves),
                           (2 * sizeof(double) + sizeof(int)) * (size_t)nb,
                           c.stream, n, nb, c.perm, c.xw, c.ytc, c.yts, c.b, c.bext, c.farx_mom, c.farx_mom1,
                           c.vfar);
        SD_HIP(hipGetLastError());
    }
    if ((c.timing & kTimePair) && (rc = timer_begin(c, c.t_pair))) return rc;
    if (c.p.field_mode == SDDDST_FIELD_EXACT)
        hipLaunchKernelGGL(k_pair_tiles<SDDDST_FIELD_EXACT>, dim3((unsigned)units), dim3(kTile), 0, c.stream, n, nb,
                           (int)nfull, c.x[slot], c.y, c.b, c.perm, c.partial, c.minr2,
                           pairtile::tolerance_r2(minPrecSqr), pairtile::FarData());
    else
        hipLaunchKernelGGL(k_pair_tiles<SDDDST_FIELD_FAST>, dim3((unsigned)units), dim3(kTile), 0, c.stream, n, nb,
                           (int)nfull, c.x[slot], c.y, c.b, c.perm, c.partial, c.minr2,
                           pairtile::tolerance_r2(minPrecSqr), fd);
    SD_HIP(hipGetLastError());
    if ((c.timing & kTimePair) && (rc = timer_end(c, c.t_pair))) return rc;
    hipLaunchKernelGGL(k_speeds_finish, dim3(blocks), dim3(threads), 0, c.stream, n, nb, c.pc, c.perm, c.partial,
                       c.x[slot],
                       c.y, c.b, c.px, c.py, c.p.A, c.p.KASQR, sigma, c.minr2, c.tol, minPrecSqr, update_tol,
                       fd.expand ? c.vfar : nullptr, c.v[out_buf]);
    SD_HIP(hipGetLastError());
    if ((c.timing & kTimeSpeeds) && (rc = timer_end(c, c.t_speeds))) return rc;
    return SDDDST_OK;
}

// ---------------------------------------------------------------------------
// Pair speeds without allocation (sdddst_speeds_noalloc).  The context path
// above keeps its partial sums, order and far-pair factors in buffers the
// context allocates; this variant keeps them in one file-scope device array,
// `scratch`, sized for at most kNoAllocMaxN dislocations, so a call allocates
// nothing.  Host code addresses the array through the pointer that
// hipGetSymbolAddress returns for the current device (taken on the first
// call on each device), never through its name.  Tiles follow index order
// (no curve sort, whose temporary storage would be one more buffer): the
// same pair terms, another fixed summation order.
namespace {

constexpr int kNoAllocMaxN = 8192;
constexpr int kNoAllocMaxBlocks = kNoAllocMaxN / kTile;
// layout of scratch: partial sums [nb][n], then six per-dislocation FP64
// arrays (cos 2 pi y, sin 2 pi y, wrapped y, wrapped x, e^{+-2 pi w}), the
// tolerance candidates (u64) and the order (int)
constexpr size_t kNoAllocPartialBytes = sizeof(double) * (size_t)kNoAllocMaxBlocks * kNoAllocMaxN;
constexpr size_t kNoAllocVecBytes = sizeof(double) * (size_t)kNoAllocMaxN;
constexpr size_t kNoAllocBytes = kNoAllocPartialBytes + 7 * kNoAllocVecBytes + sizeof(int) * (size_t)kNoAllocMaxN;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices];   // device address of scratch, per device

__global__ void k_iota(int n, int *__restrict__ perm)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) perm[i] = i;
}

int noalloc_scratch(unsigned char **out)
{
    int dev = 0;
    SD_HIP(hipGetDevice(&dev));
    SD_CHECK(dev >= 0 && dev < kMaxDevices, SDDDST_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        SD_CHECK(e == hipSuccess && p, SDDDST_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                 hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[dev];
    return SDDDST_OK;
}

}  // namespace

int speeds_noalloc(const double *x, const double *y, const double *b, int n, double sigma, int field_mode,
                   double *v, hipStream_t stream)
{
    SD_CHECK(n >= 2 && n <= kNoAllocMaxN, SDDDST_ERR_ARG,
             "sdddst_speeds_noalloc: n = %d outside [2, %d], the size of its static scratch", n, kNoAllocMaxN);
    SD_CHECK(x && y && b && v, SDDDST_ERR_ARG, "sdddst_speeds_noalloc: NULL array");
    SD_CHECK(field_mode == SDDDST_FIELD_FAST || field_mode == SDDDST_FIELD_EXACT, SDDDST_ERR_ARG,
             "sdddst_speeds_noalloc: field mode %d (fast or exact)", field_mode);
    unsigned char *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    double *partial = reinterpret_cast<double *>(base);
    double *vec = reinterpret_cast<double *>(base + kNoAllocPartialBytes);
    double *ytc = vec, *yts = vec + kNoAllocMaxN, *yw = vec + 2 * kNoAllocMaxN, *xw = vec + 3 * kNoAllocMaxN,
           *xep = vec + 4 * kNoAllocMaxN, *xem = vec + 5 * kNoAllocMaxN;
    unsigned long long *minr2 = reinterpret_cast<unsigned long long *>(vec + 6 * kNoAllocMaxN);
    int *perm = reinterpret_cast<int *>(base + kNoAllocPartialBytes + 7 * kNoAllocVecBytes);
    const int nb = (n + kTile - 1) / kTile;
    const long nfull = (long)nb * (nb - 1) / 2;
    const long units = nfull + (nb + 1) / 2;
    const int blocks = (n + 255) / 256;
    // no tolerance update: no pair is a candidate (r2_keep = 0), so every
    // pair at a far circular x distance may take the far form
    const bool far = field_mode == SDDDST_FIELD_FAST;
    hipLaunchKernelGGL(k_iota, dim3(blocks), dim3(256), 0, stream, n, perm);
    if (far) hipLaunchKernelGGL(k_ytrig, dim3(blocks), dim3(256), 0, stream, n, y, ytc, yts, yw);
    hipLaunchKernelGGL(k_speeds_prep, dim3(blocks), dim3(256), 0, stream, n, minr2, x, far ? xw : nullptr, xep, xem);
    SD_HIP(hipGetLastError());
    if (far) {
        pairtile::FarData fd;
        fd.ca = ytc;
        fd.sa = yts;
        fd.w = xw;
        fd.ep = xep;
        fd.em = xem;
        fd.yw = yw;
        hipLaunchKernelGGL(k_pair_tiles<SDDDST_FIELD_FAST>, dim3((unsigned)units), dim3(kTile), 0, stream, n, nb,
                           (int)nfull, x, y, b, perm, partial, minr2, 0.0, fd);
    } else {
        hipLaunchKernelGGL(k_pair_tiles<SDDDST_FIELD_EXACT>, dim3((unsigned)units), dim3(kTile), 0, stream, n, nb,
                           (int)nfull, x, y, b, perm, partial, minr2, 0.0, pairtile::FarData());
    }
    SD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_speeds_finish, dim3(blocks), dim3(256), 0, stream, n, nb, 0, perm, partial, x, y, b,
                       nullptr, nullptr, 0.0, 0.0, sigma, minr2, nullptr, 0.0, 0, nullptr, v);
    SD_HIP(hipGetLastError());
    return SDDDST_OK;
}

}  // namespace sdamd

extern "C" int sdddst_speeds_noalloc(const double *x, const double *y, const double *b, int n, double sigma,
                                     int field_mode, double *v, void *stream)
{
    return sdamd::speeds_noalloc(x, y, b, n, sigma, field_mode, v, static_cast<hipStream_t>(stream));
}
