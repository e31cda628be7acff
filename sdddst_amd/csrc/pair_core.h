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
//
// With Perm, tile positions are sorted positions: perm[p] is the dislocation
// at position p (the single-system kernel orders them along a Hilbert curve
// over the unit cell, so that tiles are compact in x and y and most hold
// only far pairs, whose wavefronts take the short far form throughout).
// Every pair term is still evaluated lower original index first -- the
// difference is negated when the source has the lower index, and the term
// then enters the row with a minus sign and the column with a plus sign.  The partial sums stay at sorted positions
// (coalesced stores); the finishing pass maps them back through perm.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/sdddst_amd.h"
#include "field.h"

namespace sdamd {
namespace pairtile {

constexpr int kTile = 64;
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

// The source tile in LDS, declared once by the calling kernel (one copy per
// workgroup however many tile variants the kernel instantiates).
struct TileLds {
    double x[kTile], y[kTile], b[kTile];
    // far-pair factors of the source slots: cos, sin of 2 pi y, wrapped x
    // and e^{+-2 pi x}
    double ca[kTile], sa[kTile], w[kTile], ep[kTile], em[kTile];
    int id[kTile];
};

// Per-dislocation far-pair factors of the fast field (null members: every
// pair takes the reference-form path), and optionally the extents of each
// 64-block of tile positions (block_extents), from which a tile classifies
// without wave reductions.
struct FarData {
    const double *ca = nullptr, *sa = nullptr;   // cos, sin of 2 pi y
    const double *w = nullptr, *ep = nullptr, *em = nullptr;   // x - rint(x), e^{+-2 pi w}
    const double *yw = nullptr;                  // y - rint(y)
    const double *ext = nullptr;                 // [nb][kExtWords]
    int expand = 0;                              // x-separated tiles summed elsewhere (far_expand.h)
};

// Extents of one block of tile positions: the wrapped x and y of its
// members moved by whole periods to within 1/2 of its first member's
// (min, max of each), and the raw wrapped x (min, max).
constexpr int kExtWords = 6;

// A pair takes the far form when cosh(2 pi dx) - cos(2 pi dy), as the far
// form computes it, is at least kFarMinGap: few-ulp errors in the
// exponentials and the cosine then move the term by at most ~2e-14 of
// |tau| + 1.  Tiles classify from their members' extents (below): all pairs
// far when a lower bound of the gap over the tile clears kFarMinGap, none
// when an upper bound stays below it.
//
// How a tile's pairs are evaluated: every pair in the reference form (no
// far-pair factors: the exact and ordered fields, the ensemble), every pair
// in the far form (all provably far by sdfield::gap_kind), or per pair by
// sdfield::mixed_pair (every other tile of the fast field): the far form for
// a far pair, for a near one the reference's minimum image plus the far
// images from the same factors.
//
// A far tile whose x-differences s = w_i - w_j all lie strictly inside
// (0, 1/2) (kFarPos) or all inside (-1/2, 0) (kFarNeg) -- no wrap, one
// sign: most far tiles of the curve order -- has dx = s, and the prelude's
// e^{2 pi |dx|}, e^{-2 pi |dx|} are the products e^{+-2 pi w_i} e^{-+2 pi w_j}
// it would select, formed directly: the same values bit for bit.  Inside
// (1/2, 1) (kFarWrapPos: dx = s - 1) or (-1, -1/2) (kFarWrapNeg: dx = s + 1)
// the prelude's choice is the mirror product times e^{+-2 pi}, likewise.
enum TileKind { kNearTile = 0, kFarTile = 1, kMixedTile = 2, kFarPos = 3, kFarNeg = 4, kFarWrapPos = 5,
                kFarWrapNeg = 6 };

__device__ __forceinline__ double rotate_down1(double v, int lane)
{
    // lane l receives the value of lane (l + 1) mod 64
    return __shfl(v, (lane + 1) & 63, 64);
}

// sct / e2t: the cos and exp tables of the fast field, or a copy in LDS
template <int Mode>
__device__ __forceinline__ double field_tau(double dx, double dy, const double *__restrict__ sct,
                                            const uint64_t *__restrict__ e2t)
{
    if constexpr (Mode == SDDDST_FIELD_FAST) return sdfield::xy_fast(dx, dy, sct, e2t);
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

// One pair's field for target (xi, yi) of original index io and source slot
// m, lower original index first, with the sign of the row contribution and
// without the Burgers vectors (the sweep folds b_j into the row sum and b_i
// into the column sum, and multiplies each sum by its own b once).  The
// fast field is odd bit for bit (tau(-d) = -tau(d): cos is even, the near
// image's expression and the far images' sg(dx) F(|dx|) flip sign exactly),
// and the wrap satisfies normalize(-d) = -normalize(d) except at |d| = 1/2
// exactly, so for it the target-minus-source term already equals the
// lower-first term with its sign; dx = +-0 differs only by the far images'
// rounding residue (below 1e-17).
template <int Mode, bool Perm>
__device__ __forceinline__ double pair_term(double xi, double yi, int io, int jo, double sxm, double sym,
                                            bool lower_by_lane, double &r2, const double *__restrict__ sct,
                                            const uint64_t *__restrict__ e2t)
{
    const bool lower = (Perm && Mode != SDDDST_FIELD_FAST) ? io < jo : (Perm ? true : lower_by_lane);
    const double d = xi - sxm, e = yi - sym;
    double dx, dy;
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        // the wrap as d - rint(d): the reference loop's value except at some
        // exact half-integers (+0.5 for -0.5) and the sign of a zero
        dx = sdfield::wrap_fast(lower ? d : -d);
        dy = sdfield::wrap_fast(lower ? e : -e);
    } else {
        dx = sdfield::normalize(lower ? d : -d);
        dy = sdfield::normalize(lower ? e : -e);
    }
    const double t = field_tau<Mode>(dx, dy, sct, e2t);
    r2 = dx * dx + dy * dy;
    return lower ? t : -t;
}

// The far-pair factors of one dislocation (a target lane's registers).
struct FarSelf {
    double ca = 0.0, sa = 0.0, w = 0.0, ep = 0.0, em = 0.0;
};

__device__ __forceinline__ FarSelf far_load(const FarData &f, int i)
{
    FarSelf r;
    r.ca = f.ca[i];
    r.sa = f.sa[i];
    r.w = f.w[i];
    r.ep = f.ep[i];
    r.em = f.em[i];
    return r;
}

__device__ __forceinline__ void far_stage(TileLds &L, int l, const FarSelf &v)
{
    L.ca[l] = v.ca;
    L.sa[l] = v.sa;
    L.w[l] = v.w;
    L.ep[l] = v.ep;
    L.em[l] = v.em;
}

// min and max over the 64 lanes of a wrapped coordinate v (NaN outside the
// system) after moving each lane's v by a whole period to within 1/2 of lane
// 0's (a lane inside the system in every tile), so that a tile straddling
// the cell's edge keeps a short interval.  The shifted values are exact up
// to one rounding of v -+ 1 (below 2^-53).
__device__ __forceinline__ void unwrapped_extent(double v, double &vmin, double &vmax)
{
    const double r = __shfl(v, 0, 64);
    const double u = v - __builtin_rint(v - r);
    vmin = u;
    vmax = u;
    for (int off = 32; off > 0; off >>= 1) {
        vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    }
}

// The extents of a block from its members' wrapped x and y (NaN outside the
// system; the first member is inside): called by one whole 64-lane
// wavefront, the result on every lane, in the order of kExtWords.
__device__ __forceinline__ void block_extents(double w, double yw, double out[kExtWords])
{
    unwrapped_extent(w, out[0], out[1]);
    unwrapped_extent(yw, out[2], out[3]);
    double lo = w, hi = w;
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    out[4] = lo;
    out[5] = hi;
}

// The tile kind from the wrapped x and y of the target lanes (tw, ty) and
// the source slots (sw, sy); the same on every lane.
__device__ __forceinline__ int tile_kind(double tw, double sw, double ty, double sy)
{
    double tmin, tmax, smin, smax, xmin, xmax, ymin, ymax;
    unwrapped_extent(tw, tmin, tmax);
    unwrapped_extent(sw, smin, smax);
    sdfield::circ_range(tmin, tmax, smin, smax, xmin, xmax);
    unwrapped_extent(ty, tmin, tmax);
    unwrapped_extent(sy, smin, smax);
    sdfield::circ_range(tmin, tmax, smin, smax, ymin, ymax);
    return __builtin_amdgcn_readfirstlane(sdfield::gap_kind(xmin, xmax, ymin, ymax));
}


// The interval that holds every x-difference s = w_i - w_j of a far tile,
// with a margin of 1e-12 for the rounding of s: (0, 1/2) -> kFarPos,
// (-1/2, 0) -> kFarNeg, (1/2, 1) -> kFarWrapPos, (-1, -1/2) -> kFarWrapNeg,
// none of them -> kFarTile; from the extremes of the wrapped x of the
// targets (tmin, tmax) and of the sources (smin, smax).  The same on every
// lane.
__device__ __forceinline__ int far_kind_from(double tmin, double tmax, double smin, double smax)
{
    const double lo = tmin - smax, hi = tmax - smin, e = 1e-12;
    int k = kFarTile;
    if (lo > e && hi < 0.5 - e) k = kFarPos;
    else if (hi < -e && lo > -0.5 + e) k = kFarNeg;
    else if (lo > 0.5 + e && hi < 1.0 - e) k = kFarWrapPos;
    else if (hi < -0.5 - e && lo > -1.0 + e) k = kFarWrapNeg;
    return __builtin_amdgcn_readfirstlane(k);
}

// far_kind_from with the extremes taken over the wrapped x of the target
// lanes (tw) and the source slots (sw), NaN outside the system (lane 0 is
// inside in every tile).
__device__ __forceinline__ int far_tile_kind(double tw, double sw)
{
    double tmin = tw, tmax = tw, smin = sw, smax = sw;
    for (int off = 32; off > 0; off >>= 1) {
        tmin = fmin(tmin, __shfl_xor(tmin, off, 64));
        tmax = fmax(tmax, __shfl_xor(tmax, off, 64));
        smin = fmin(smin, __shfl_xor(smin, off, 64));
        smax = fmax(smax, __shfl_xor(smax, off, 64));
    }
    return far_kind_from(tmin, tmax, smin, smax);
}


// tile_kind, and far_tile_kind for an all-far tile, from the precomputed
// extents of the target block (et) and the source block (es): the same
// intervals the reductions form, so the same kind.
__device__ __forceinline__ int tile_kind_ext(const double *__restrict__ et, const double *__restrict__ es)
{
    double xmin, xmax, ymin, ymax;
    sdfield::circ_range(et[0], et[1], es[0], es[1], xmin, xmax);
    sdfield::circ_range(et[2], et[3], es[2], es[3], ymin, ymax);
    const int kind = __builtin_amdgcn_readfirstlane(sdfield::gap_kind(xmin, xmax, ymin, ymax));
    return kind == kFarTile ? far_kind_from(et[4], et[5], es[4], es[5]) : kind;
}

// One pair of a far or mixed tile: the far form when its gap
// cosh(2 pi dx) - cos(2 pi dy) (far_gap) is at least kFarMinGap (always in a
// far tile), else the reference's minimum image plus the far images from the
// factors (sdfield::mixed_pair).  The choice is symmetric in the two
// dislocations bit for bit (the prelude's t and ti and cos_diff's commuted
// products are).  r2 is set for near pairs only (far pairs are never
// tolerance candidates: their r^2 > kFarMinR2 exceeds every r2_keep the
// caller passes on this path).
template <int Mode, bool Perm, int Kind>
__device__ __forceinline__ double tile_term(double xi, double yi, int io, int jm, const TileLds &L, int m,
                                            const FarSelf &fi, bool lower_by_lane, double &r2,
                                            const double *__restrict__ sct, const uint64_t *__restrict__ e2t)
{
    if constexpr (Kind == kNearTile) {
        return pair_term<Mode, Perm>(xi, yi, io, jm, L.x[m], L.y[m], lower_by_lane, r2, sct, e2t);
    } else if constexpr (Kind == kFarPos || Kind == kFarNeg) {
        // t = e^{2 pi |s|}: e^{2 pi w_i} e^{-2 pi w_j} for s > 0, the mirror
        // product for s < 0 (the far prelude's choices, same operands)
        const double cy = sdfield::cos_diff(fi.ca, fi.sa, L.ca[m], L.sa[m]);
        const double u = fi.ep * L.em[m], v = fi.em * L.ep[m];
        r2 = 1.0;
        return Kind == kFarPos ? sdfield::xy_far(fi.w - L.w[m], cy, u, v)
                               : sdfield::xy_far(fi.w - L.w[m], cy, v, u);
    } else if constexpr (Kind == kFarWrapPos || Kind == kFarWrapNeg) {
        // dx = s -+ 1: t = e^{2 pi (1 -+ s)} is the mirror product times
        // e^{2 pi}, ti the product times e^{-2 pi} (as the prelude forms them)
        const double cy = sdfield::cos_diff(fi.ca, fi.sa, L.ca[m], L.sa[m]);
        const double u = fi.ep * L.em[m], v = fi.em * L.ep[m];
        const double s = fi.w - L.w[m];
        r2 = 1.0;
        return Kind == kFarWrapPos
                   ? sdfield::xy_far(s - 1.0, cy, v * sdlibm_exp_2pik[1], u * sdlibm_exp_m2pik[1])
                   : sdfield::xy_far(s + 1.0, cy, u * sdlibm_exp_2pik[1], v * sdlibm_exp_m2pik[1]);
    } else {
        const double cy = sdfield::cos_diff(fi.ca, fi.sa, L.ca[m], L.sa[m]);
        const sdfield::FarPrelude f = sdfield::far_prelude(fi.w, L.w[m], fi.ep, fi.em, L.ep[m], L.em[m]);
        if constexpr (Kind == kFarTile) {
            r2 = 1.0;
            return sdfield::xy_far(f.dx, cy, f.t, f.ti);
        } else {
            // the fast field only (the far-pair factors exist for it alone):
            // odd bit for bit, so target minus source serves every pair
            static_assert(Mode == SDDDST_FIELD_FAST, "mixed tiles are a fast-field form");
            bool far;
            return sdfield::mixed_pair(f, cy, sdfield::wrap_fast(xi - L.x[m]), sdfield::wrap_fast(yi - L.y[m]), far,
                                       r2, sct, e2t);
        }
    }
}

// Called by one whole 64-lane workgroup.  Full: both blocks lie inside n
// and bi < bj, so every pair of the tile counts.  sct / e2t (fast field
// only) may point at LDS copies of the tables written before the call: the
// tile's barrier orders them before the first read.  With far-pair factors
// (fast field) a pair whose gap cosh(2 pi dx) - cos(2 pi dy) is at least
// kFarMinGap takes the far form (cos by angle addition, the exponentials from
// per-dislocation factors); the choice is made per pair, so a pair's value
// does not depend on the tile that holds it, and the tile's x and y extents
// decide once whether all, none or some of its pairs are far.
template <int Mode, bool Perm = false, bool Full = false>
__device__ __forceinline__ void tile(TileLds &L, int bi, int bj, int n, const double *__restrict__ x,
                                     const double *__restrict__ y, const double *__restrict__ b,
                                     const int *__restrict__ perm, double *__restrict__ partial,
                                     unsigned long long *__restrict__ minr2, double r2_keep,
                                     const double *__restrict__ sct = sdlibm_sincos_tab,
                                     const uint64_t *__restrict__ e2t = sdlibm_exp2_tab, FarData fd = FarData())
{
    double *sx = L.x, *sy = L.y, *sb = L.b;
    int *sid = L.id;
    const int l = (int)threadIdx.x;
    const int p = bi * kTile + l;              // target position
    const int q = bj * kTile + l;              // source position held in slot l
    const int io = Perm ? (p < n ? perm[p] : 0) : p;
    const int jo = Perm ? (q < n ? perm[q] : 0) : q;
    sx[l] = q < n ? x[jo] : 0.0;
    sy[l] = q < n ? y[jo] : 0.0;
    sb[l] = q < n ? b[jo] : 0.0;
    sid[l] = jo;
    const double xi = p < n ? x[io] : 0.0;
    const double yi = p < n ? y[io] : 0.0;
    const double bI = p < n ? b[io] : 0.0;
    int kind = kNearTile;
    FarSelf fi;
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        if (fd.w) {
            const double nan = __builtin_nan("");
            const FarSelf src = q < n ? far_load(fd, jo) : FarSelf();
            far_stage(L, l, src);
            if (p < n) fi = far_load(fd, io);
            if (fd.ext) {
                kind = tile_kind_ext(fd.ext + (size_t)bi * kExtWords, fd.ext + (size_t)bj * kExtWords);
            } else {
                kind = tile_kind(p < n ? fi.w : nan, q < n ? src.w : nan, p < n ? fd.yw[io] : nan,
                                 q < n ? fd.yw[jo] : nan);
                if (kind == kFarTile) kind = far_tile_kind(p < n ? fi.w : nan, q < n ? src.w : nan);
            }
            // every pair of a tile that is not all far takes the mixed form
            // (so a near pair's value does not depend on its tile's kind)
            if (kind == kNearTile) kind = kMixedTile;
        }
    }
    __syncthreads();
    const bool diag = (bi == bj);
    auto sweep = [&](auto kind_tag) {
        constexpr int K = decltype(kind_tag)::value;
        double row = 0.0, col = 0.0;
        for (int s = 0; s < kTile; ++s) {
            const int m = (l + s) & 63;
            const int j = bj * kTile + m;
            const bool valid = Full || ((p < n) && (j < n) && (!diag || l < m));
            // the source's original index: for the lower-first rule (exact
            // field) and for the rare tolerance candidates
            const int jm = (Perm && Mode != SDDDST_FIELD_FAST) ? sid[m] : j;
            double r2;
            // without Perm the target has the lower index in every tile it meets
            const double u = tile_term<Mode, Perm, K>(xi, yi, io, jm, L, m, fi, true, r2, sct, e2t);
            if (valid) {
                // term b_i b_j u: b_j here, b_i on the row sum below; the
                // column sum of source j gets -b_i u here and b_j below
                row = __builtin_fma(sb[m], u, row);
                col = __builtin_fma(-bI, u, col);
                if ((K == kNearTile || K == kMixedTile) && r2 < r2_keep) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong(r2);
                    atomicMin(&minr2[io], bits);
                    atomicMin(&minr2[Perm ? sid[m] : j], bits);
                }
            }
            col = rotate_down1(col, l);
        }
        // after the 64 sweeps lane l holds the column sum of source slot l
        if (diag) {
            if (p < n) partial[(size_t)bi * n + p] = bI * (row + col);
        } else {
            if (p < n) partial[(size_t)bj * n + p] = bI * row;
            if (q < n) partial[(size_t)bi * n + q] = sb[l] * col;
        }
    };
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        if (kind == kFarPos) sweep(std::integral_constant<int, kFarPos>{});
        else if (kind == kFarNeg) sweep(std::integral_constant<int, kFarNeg>{});
        else if (kind == kFarWrapPos) sweep(std::integral_constant<int, kFarWrapPos>{});
        else if (kind == kFarWrapNeg) sweep(std::integral_constant<int, kFarWrapNeg>{});
        else if (kind == kFarTile) sweep(std::integral_constant<int, kFarTile>{});
        else if (kind == kMixedTile) sweep(std::integral_constant<int, kMixedTile>{});
        else sweep(std::integral_constant<int, kNearTile>{});
    } else {
        sweep(std::integral_constant<int, kNearTile>{});
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
// as in the off-diagonal tiles: when the source has the lower index the
// difference is negated and the term enters the row with a minus sign and
// the column with a plus sign.  Called by one whole 64-lane workgroup.
template <int Mode, bool Perm = false>
__device__ __forceinline__ void diag_tile(TileLds &L, int bk, int n, const double *__restrict__ x,
                                          const double *__restrict__ y,
                                          const double *__restrict__ b, const int *__restrict__ perm,
                                          double *__restrict__ partial, unsigned long long *__restrict__ minr2,
                                          double r2_keep, const double *__restrict__ sct = sdlibm_sincos_tab,
                                          const uint64_t *__restrict__ e2t = sdlibm_exp2_tab, FarData fd = FarData())
{
    double *sx = L.x, *sy = L.y, *sb = L.b;
    int *sid = L.id;
    const int l = (int)threadIdx.x;
    const int p = bk * kTile + l;
    const int io = Perm ? (p < n ? perm[p] : 0) : p;
    const double xi = p < n ? x[io] : 0.0;
    const double yi = p < n ? y[io] : 0.0;
    const double bI = p < n ? b[io] : 0.0;
    sx[l] = xi;
    sy[l] = yi;
    sb[l] = bI;
    sid[l] = io;
    // a block against itself is never all far (its intervals overlap
    // themselves): with far-pair factors every pair takes the mixed form, as
    // in the off-diagonal tiles that are not all far
    int kind = kNearTile;
    FarSelf fi;
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        if (fd.w) {
            kind = kMixedTile;
            if (p < n) fi = far_load(fd, io);
            far_stage(L, l, fi);
        }
    }
    __syncthreads();
    auto sweep = [&](auto kind_tag) {
        constexpr int K = decltype(kind_tag)::value;
        double row = 0.0, col = 0.0;
        for (int s = 1; s <= kTile / 2; ++s) {
            const int m = (l + s) & 63;
            const int j = bk * kTile + m;
            const bool valid = (p < n) && (j < n) && (s < kTile / 2 || l < kTile / 2);
            const int jm = (Perm && Mode != SDDDST_FIELD_FAST) ? sid[m] : j;
            double r2;
            const double u = tile_term<Mode, Perm, K>(xi, yi, io, jm, L, m, fi, m > l, r2, sct, e2t);
            if (valid) {
                row = __builtin_fma(sb[m], u, row);
                col = __builtin_fma(-bI, u, col);
                if (r2 < r2_keep) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong(r2);
                    atomicMin(&minr2[io], bits);
                    atomicMin(&minr2[Perm ? sid[m] : j], bits);
                }
            }
            col = rotate_down1(col, l);
        }
        // after sweeps 1..32 the column sum of source m sits in lane (m - 33) mod 64
        const double cs = __shfl(col, (l + 31) & 63, 64);
        if (p < n) partial[(size_t)bk * n + p] = bI * (row + cs);
    };
    if constexpr (Mode == SDDDST_FIELD_FAST) {
        if (kind == kMixedTile) sweep(std::integral_constant<int, kMixedTile>{});
        else sweep(std::integral_constant<int, kNearTile>{});
    } else {
        sweep(std::integral_constant<int, kNearTile>{});
    }
}

// r^2 below which a pair can change a tolerance: 400 prec^2 with a margin
// for the roundings of r^2 * 0.25 * 1e-2.
__host__ __device__ inline double tolerance_r2(double minPrecSqr) { return minPrecSqr * 400.0 * (1.0 + 1e-6); }

}  // namespace pairtile
}  // namespace sdamd
