// Far field of x-separated tiles by the Fourier series of the field in y.
//
// For |dx| in (0, 1/2) the fast field's sum of the reference's images
// (AnalyticField.cpp:145-160: the images dx + k, k = -3..3, then
// (1 - dx) f(dx + 4), f(dx - 4) and dx f(dx - 5) for dx > 0, the mirror set
// for dx < 0) takes, image by image, the Fourier series
//     f(u) = 2 pi^2 u (C cy - 1) / (C - cy)^2 = 4 pi^2 u sum_n n q^n cos(2 pi n dy),
// C = cosh(2 pi u), q = e^{-2 pi |u|}.  Collected by n, for dx > 0,
//     tau(dx, dy) = 4 pi^2 sum_n n cos(2 pi n dy)
//                   [e^{-2 pi n dx} alpha_n(dx) + e^{-2 pi n (1 - dx)} beta_n(dx)],
//     alpha_n(dx) = -g4 dx^2 + (1 + g1 + g2 + g3 - 3 g4) dx + (g1 + 2 g2 + 3 g3 + 4 g4),
//     beta_n(dx)  =  g4 dx^2 + (1 + g1 + g2 + g3 - 5 g4) dx - (1 + 2 g1 + 3 g2 + 4 g3),
// g_k = e^{-2 pi n k} (beta_n carries the factor e^{-2 pi n} of its nearest
// image in the exponential), and tau(-dx, dy) = -tau(dx, dy).
//
// With dx = D + xi_i - xi_j (xi the offset from the block's x-centre, D the
// centres' difference plus the wrap offset) and cos(2 pi n dy) =
// Re e^{2 pi i n y_i} e^{-2 pi i n y_j}, every term separates into a factor
// of the target, one of the tile and one of the source: a source block
// keeps the moments sum_j b_j xi_j^p e^{+-2 pi n xi_j} e^{-2 pi i n y_j}
// (p = 0, 1; p = 2 for n = 1, where the g4 terms still count), a target
// block gathers, over every source block of its x-separated tiles, three
// coefficients per n and family (powers 0..2 of xi_i), and each target sums
// n Re(psi^n (L0 + xi_i (L1 + xi_i L2))) with psi = e^{-+2 pi xi_i}
// e^{2 pi i y_i}.  A tile qualifies when every x-difference s = w_i - w_j
// (w = x - rint(x)) of its pairs lies on one side of 0 and of a wrap, with
// |dx| >= kMinDx: the first kTerms terms then leave less than 1e-17 of the
// field (q <= e^{-2 pi kMinDx}), kTermsB those of beta (q <= e^{-pi}).
// Measured against the reference's sums over such tiles the result is
// within 1e-15 of (sum |tau| + 1); the pair kernel skips these tiles.
#pragma once

#include <hip/hip_runtime.h>

#include "field.h"

namespace sdamd {
namespace farx {

constexpr double kMinDx = 0.11;
constexpr int kTerms = 64;    // n = 1..kTerms (one lane each)
constexpr int kTermsB = 14;   // the beta terms, n = 1..kTermsB
// per block and term n: e^{2 pi n X}, e^{-2 pi n X}, then the complex
// moments M0+, M1+, M0-, M1- (xi^p e^{+-2 pi n xi} b e^{-2 pi i n y})
constexpr int kWords = 10;
// per block, n = 1 only: M2+, M2- (complex)
constexpr int kWords1 = 4;

// The x-centre of a block from its raw wrapped-x extent.
__device__ __forceinline__ double centre(const double *__restrict__ ext) { return 0.5 * (ext[4] + ext[5]); }

// Whether the tile of target block t and source block s qualifies, from the
// blocks' raw wrapped-x extents (minimum and maximum; NaN-free for every
// block with a member): 0 no; 1 s in [kMinDx, 1/2) (dx = s); 2 s in (-1/2,
// -kMinDx] (dx = s); 3 s in (1/2, 1 - kMinDx] (dx = s - 1); 4 s in
// [-1 + kMinDx, -1/2) (dx = s + 1).  The margins at the half and whole
// periods are those of the pair kernel's far kinds.
__device__ __forceinline__ int tile_code_x(double tmin, double tmax, double smin, double smax)
{
    const double lo = tmin - smax, hi = tmax - smin, e = 1e-12;
    int k = 0;
    if (lo >= kMinDx && hi < 0.5 - e) k = 1;
    else if (hi <= -kMinDx && lo > -0.5 + e) k = 2;
    else if (lo > 0.5 + e && hi <= 1.0 - kMinDx) k = 3;
    else if (hi < -0.5 - e && lo >= -1.0 + kMinDx) k = 4;
    return k;
}

// The same from the blocks' extent records (words 4 and 5), for a tile the
// whole wavefront holds.
__device__ __forceinline__ int tile_code(const double *__restrict__ et, const double *__restrict__ es)
{
    return __builtin_amdgcn_readfirstlane(tile_code_x(et[4], et[5], es[4], es[5]));
}

// Side (+1: dx > 0) and wrap offset o (dx = w_i - w_j + o) of the pairs of
// a qualifying tile, seen from its target block; from its source block both
// change sign.
__device__ __forceinline__ void code_side(int code, int &sigma, double &o)
{
    sigma = (code == 1 || code == 4) ? 1 : -1;
    o = code == 3 ? -1.0 : (code == 4 ? 1.0 : 0.0);
}

struct Cplx {
    double re, im;
};
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b)
{
    return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cplx cscale(Cplx a, double s) { return Cplx{a.re * s, a.im * s}; }
__device__ __forceinline__ Cplx cadd(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cplx cpow(Cplx a, int n)
{
    Cplx r{1.0, 0.0};
    while (n > 0) {
        if (n & 1) r = cmul(r, a);
        a = cmul(a, a);
        n >>= 1;
    }
    return r;
}

// Polynomial coefficients of term n: alpha = (a2, a1, a0), beta = (b2, b1,
// b0); the squares only for n = 1 (e^{-16 pi} below).
struct Poly {
    double q2, q1, q0;
};
__device__ __forceinline__ void term_polys(int n, Poly &alpha, Poly &beta)
{
    const double g1 = exp(-sdfield::kTwoPi * n);
    const double g2 = g1 * g1, g3 = g2 * g1, g4 = n == 1 ? g2 * g2 : 0.0;
    alpha = Poly{-g4, 1.0 + g1 + g2 + g3 - 3.0 * g4, g1 + 2.0 * g2 + 3.0 * g3 + 4.0 * g4};
    beta = Poly{g4, 1.0 + g1 + g2 + g3 - 5.0 * g4, -(1.0 + 2.0 * g1 + 3.0 * g2 + 4.0 * g3)};
}

// L[0..2] += e * (bracket in powers of xi_i) for the polynomial argument
// p + sj xi_j with p = P + si xi_i:
//   poly(p + sj xi_j) = c0(p) + sj c1(p) xi_j + q2 xi_j^2,
//   c0(P + si xi_i) = c0(P) + si c1(P) xi_i + q2 xi_i^2,
//   c1(P + si xi_i) = c1(P) + 2 q2 si xi_i.
__device__ __forceinline__ void gather(Cplx &L0, Cplx &L1, Cplx &L2, double e, double P, const Poly &q, double si,
                                       double sj, Cplx M0, Cplx M1, Cplx M2)
{
    const double c0 = (q.q2 * P + q.q1) * P + q.q0;
    const double c1 = 2.0 * q.q2 * P + q.q1;
    L0 = cadd(L0, cscale(cadd(cadd(cscale(M0, c0), cscale(M1, sj * c1)), cscale(M2, q.q2)), e));
    L1 = cadd(L1, cscale(cadd(cscale(M0, si * c1), cscale(M1, 2.0 * q.q2 * si * sj)), e));
    L2 = cadd(L2, cscale(M0, q.q2 * e));
}

}  // namespace farx
}  // namespace sdamd
