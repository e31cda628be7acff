// The analytic periodic shear-stress field of an edge dislocation in the unit
// cell [-0.5, 0.5)^2 (reference: src/Fields/AnalyticField.cpp) and the
// periodic wrap of a coordinate difference (include/utility.h:29-40).
//
// Each function keeps the reference's expression tree and evaluation order
// so that, with the bit-compatible elementary functions of libm_compat.h and
// contraction disabled, the device reproduces the reference's per-pair
// values.  pow(v, 2.0) is written v * v: an optimised build of the reference
// folds it the same way and the two agree bit for bit.
//
// xy_fast() is the production form of xy(): the image at the
// minimum-image offset (the only one that can suffer cancellation) is
// evaluated exactly as the reference does; the nine far images take their
// cosh from one shared exp(2*pi*dx) and fixed exp(+-2*pi*k) factors.  The
// two images within 1.5 of the cell are summed as one fraction (numerator
// and product of the denominators (cosh - cos)^2; one division); the seven
// beyond it take their Fourier series in y to the order that reaches 1e-17.
// Far-image terms then differ from the reference by a few ulps of terms
// that are at least e^-pi smaller than the near term.
#pragma once

#include "libm_compat.h"

namespace sdfield {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kPiSq = kPi * kPi;       // == pow(M_PI, 2.0) as folded by the compiler
constexpr int kImages = 4;                // ANALYTIC_FIELD_N (include/constants.h:25)

// utility.h:29-40.  Each +-1 step of the reference loop is exact, so the
// loop equals d - k for the integer k that lands in [-0.5, 0.5); a zero
// (either sign) passes through the loop unchanged.  d - rint(d) is exact and
// lies in [-0.5, 0.5] (ties round to even), so only +0.5 needs the step.
SDLIBM_FN double normalize(double d)
{
    if (d == 0.0) return d;
    double r = d - __builtin_rint(d);
    if (r >= 0.5) r -= 1.0;
    return r;
}

// normalize() without its two exact-boundary adjustments, for the fast
// field: d - rint(d) lies in [-0.5, 0.5] and equals normalize(d) except at
// d = k + 1/2 with k even (+0.5 instead of -0.5; rint ties to even) and for
// d = -0 (+0).
SDLIBM_FN double wrap_fast(double d) { return d - __builtin_rint(d); }

// AnalyticField.cpp:33-38 (taylor branch of f, r^2 <= 1e-6).
SDLIBM_FN double f_taylor(double dx, double dy)
{
    const double dx2pi = dx * 2.0 * kPi;
    const double dy2pi = dy * 2.0 * kPi;
    const double x2 = dx2pi * dx2pi, x4 = sdlibm::powi(dx2pi, 4), x6 = sdlibm::powi(dx2pi, 6);
    const double y2 = dy2pi * dy2pi, y4 = sdlibm::powi(dy2pi, 4), y6 = sdlibm::powi(dy2pi, 6);
    const double a = x6 / 720.0 + x4 / 24.0 + x2 * 0.5;
    const double b = -y6 / 720.0 + y4 / 24.0 - y2 * 0.5;
    const double c = (y6 + x6) / 720. + (x4 - y4) / 24.0 + (x2 + y2) * 0.5;
    return ((a * b + b + a) * dx) / (c * c) * 2.0 * kPiSq;
}

// AnalyticField.cpp:24-39 with cosh(2 pi dx) supplied by the caller.
SDLIBM_FN double f_given_cosh(double dx, double cosh2pix, double cos2piy)
{
    return dx * (cosh2pix * cos2piy - 1.0) / ((cosh2pix - cos2piy) * (cosh2pix - cos2piy)) * 2. * kPi * kPi;
}

// AnalyticField.cpp:24-39.
SDLIBM_FN double f(double dx, double cos2piy, double dy)
{
    if (dx * dx + dy * dy > 1e-6) return f_given_cosh(dx, sdlibm::cosh(dx * 2.0 * kPi), cos2piy);
    return f_taylor(dx, dy);
}

// AnalyticField.cpp:65-90 (taylor branch of f_dx).
SDLIBM_FN double f_dx_taylor(double dx, double dy)
{
    const double X = dx * 2.0 * kPi;
    const double Y = dy * 2.0 * kPi;
    const double X2 = X * X, X3 = sdlibm::powi(X, 3), X4 = sdlibm::powi(X, 4), X5 = sdlibm::powi(X, 5),
                 X6 = sdlibm::powi(X, 6), X7 = sdlibm::powi(X, 7);
    const double Y2 = Y * Y, Y4 = sdlibm::powi(Y, 4), Y6 = sdlibm::powi(Y, 6);
    const double a = X6 / 720.0 + X4 / 24.0 + X2 * 0.5;
    const double b = -Y6 / 720.0 + Y4 / 24.0 - Y2 * 0.5;
    const double c = (Y6 + X6) / 720. + (X4 - Y4) / 24.0 + (X2 + Y2) * 0.5;
    const double cs = -(X7 * Y6) / 3628800.0 - (X5 * Y6) / 86400.0 - (X3 * Y6) / 4320.0 - (Y6 * X) / 720.0 +
                      (X7 * Y4) / 120960.0 + (X5 * Y4) / 2880.0 + (X3 * Y4) / 144.0 + (Y4 * X) / 24.0 -
                      (X7 * Y2) / 1080.0 - (X5 * Y2) / 240.0 - (X3 * Y2) / 12.0 - (Y2 * X) * 0.5 +
                      X7 / 5040.0 + X5 / 120.0 + X3 / 6.0 + X;
    const double c2 = c * c, c3 = sdlibm::powi(c, 3);
    return ((a * b + b + a) / c2 + ((cs * dx) / c2 - (a * b + b + a) / c3 * dx * sdlibm::sinh(X) * 2.0) * kPi * 2.0) *
           2.0 * kPiSq;
}

// AnalyticField.cpp:53-93.
SDLIBM_FN double f_dx(double dx, double cos2piy, double dy)
{
    const double ch = sdlibm::cosh(kPi * 2.0 * dx);
    const double sh = sdlibm::sinh(kPi * 2.0 * dx);
    if (dx * dx + dy * dy > 1e-6) {
        const double d = ch - cos2piy;
        const double d2 = d * d;
        const double d3 = sdlibm::powi(d, 3);
        return ((ch * cos2piy - 1.0) / d2 +
                dx * (sh * cos2piy * 2.0 * kPi / d2 - (ch * cos2piy - 1.0) / d3 * 4.0 * kPi * sh)) *
               kPi * kPi * 2.0;
    }
    return f_dx_taylor(dx, dy);
}

// AnalyticField.cpp:145-160, exact form.  Branch-free selection of the
// dx < 0 / dx >= 0 image sets: with s = +1 for dx < 0 and -1 otherwise the
// reference's two expressions are -s*dx*f(dx+5s) + (1+s*dx)*f(dx-4s) + f(dx+4s)
// + g3(dx), and every sign flip is exact.
SDLIBM_FN double xy(double dx, double dy)
{
    const double cy = sdlibm::cos(kPi * 2.0 * dy);
    const bool neg = dx < 0;
    const double s = neg ? 1.0 : -1.0;
    const double sdx = neg ? -dx : dx;          // -s*dx
    double g = f(dx, cy, dy);
    for (int k = 1; k < kImages; ++k) g = g + f(dx - (double)k, cy, dy) + f(dx + (double)k, cy, dy);
    return sdx * f(dx + 5.0 * s, cy, dy) + (1.0 + s * dx) * f(dx - 4.0 * s, cy, dy) + f(dx + 4.0 * s, cy, dy) + g;
}

// AnalyticField.cpp:162-181.
SDLIBM_FN double xy_diff_x(double dx, double dy)
{
    const double cy = sdlibm::cos(kPi * 2.0 * dy);
    const double N = (double)kImages;
    double g = f_dx(dx, cy, dy);
    for (int k = 1; k < kImages; ++k) g = g + f_dx(dx - (double)k, cy, dy) + f_dx(dx + (double)k, cy, dy);
    if (dx < 0) {
        const double a = dx + N + 1.0;
        return -f(a, cy, dy) - dx * f_dx(a, cy, dy) + f_dx(dx - N, cy, dy) * (1.0 + dx) + f(dx - N, cy, dy) +
               f_dx(dx + N, cy, dy) + g;
    }
    const double a = dx - N - 1.0;
    return f(a, cy, dy) + dx * f_dx(a, cy, dy) + f_dx(dx + N, cy, dy) * (1.0 - dx) - f(dx + N, cy, dy) +
           f_dx(dx - N, cy, dy) + g;
}

// cos(p) for |p| <= 2 pi (sdlibm::cos covers the first two quadrants
// glibc's reduction forms, n = round(|p| 2/pi) <= 2) in glibc's way: for
// n = 3 and 4 the argument a + da = |p| - n pi/2 (high part exact by
// Sterbenz, the constant's low part carried in da), then sin(a + da) for
// n = 3 and cos(a + da) for n = 4 by the table method.
SDLIBM_FN double cos_2pi(double p)
{
    const double ap = sdlibm::fabsd(p);
    if (sdlibm::hiword(ap) < 0x400368fdu) return sdlibm::cos(p);   // |p| < 2.426265
    const double t = ap * 0x1.45f306dc9c883p-1 + 0x1.8p52;
    const int n = (int)(t - 0x1.8p52);
    if (n <= 2) return sdlibm::cos(p);
    const double hi = n == 3 ? 0x1.2d97c7f3321d2p2 : 0x1.921fb54442d18p2;   // 3 pi/2, 2 pi
    const double lo = n == 3 ? 0x1.a79394c9e8a0ap-53 : 0x1.1a62633145c07p-52;
    const double r = ap - hi;
    const double a = r - lo;
    const double da = (r - a) - lo;
    return n == 3 ? sdlibm::detail::do_sin(a, da) : sdlibm::detail::do_cos(a, da);
}

// AnalyticField.cpp:95-106: the second x-derivative of f, as the reference
// writes it (no small-r branch, and without the factor 2 pi^2 that f and
// f_dx carry).
SDLIBM_FN double f_dx2(double dx, double cos2piy, double dy)
{
    const double ch = sdlibm::cosh(kPi * 2.0 * dx);
    const double sh = sdlibm::sinh(kPi * 2.0 * dx);
    const double c4y = cos_2pi(4.0 * kPi * dy);
    const double ch2 = sdlibm::cosh(kPi * 4.0 * dx);
    const double sh2 = sdlibm::sinh(kPi * 4.0 * dx);
    const double num = sh * cos2piy * (cos2piy * cos2piy - 2.0) + 2.0 * kPi * dx * (sh * sh) * (c4y - 2.0) -
                       kPi * dx * sdlibm::powi(ch, 3) * cos2piy +
                       0.5 * kPi * dx * ch * cos2piy * (2.0 * ch2 + c4y - 5.0) + (ch * ch) * (2.0 * kPi * dx - sh * cos2piy) +
                       sh2;
    return (4.0 * kPi * num) / sdlibm::powi(cos2piy - ch, 4);
}

// AnalyticField.cpp:183-202: the reference's d^2 tau_xy / dx^2 (images of
// f_dx2, boundary images through f_dx as well), used only by the
// eigenvalue-derivative option of the EV analysis.
SDLIBM_FN double xy_diff_x2(double dx, double dy)
{
    const double cy = sdlibm::cos(kPi * 2.0 * dy);
    const double N = (double)kImages;
    double g = f_dx2(dx, cy, dy);
    for (int k = 1; k < kImages; ++k) g = g + f_dx2(dx - (double)k, cy, dy) + f_dx2(dx + (double)k, cy, dy);
    if (dx < 0) {
        const double a = dx + N + 1.0;
        return -2.0 * f_dx(a, cy, dy) - dx * f_dx2(a, cy, dy) + f_dx2(dx - N, cy, dy) * (1.0 + dx) +
               2.0 * f_dx(dx - N, cy, dy) + f_dx2(dx + N, cy, dy) + g;
    }
    const double a = dx - N - 1.0;
    return 2.0 * f_dx(a, cy, dy) + dx * f_dx2(a, cy, dy) + f_dx2(dx + N, cy, dy) * (1.0 - dx) -
           2.0 * f_dx(dx + N, cy, dy) + f_dx2(dx - N, cy, dy) + g;
}


// cos and sin of 2 pi y for y wrapped into [-0.5, 0.5] (exactly: y - rint(y)).
SDLIBM_FN void ytrig(double y, double &c, double &s)
{
    const double a = kPi * 2.0 * (y - __builtin_rint(y));
    c = sdlibm::cos(a);
    s = sdlibm::sin(a);
}

// cos(2 pi (y_i - y_j)) from the two dislocations' ytrig values (angle
// addition; symmetric in i and j bit for bit).  Within a few ulps of the
// reference's cos(2 pi dy): used only for far pairs, those with
// cosh(2 pi dx) - cos(2 pi dy) >= kFarMinGap, where the field's sensitivity
// to it (and to few-ulp exponentials) is low.  The far form stays within
// 2.2e-14 of the reference's term relative to |tau| + 1 (3e6 samples;
// 99.99th percentile 1.3e-14, mean 3e-16) -- below the reference's own
// truncation of the image sum at +-4 images with a half-weight fifth.
SDLIBM_FN double cos_diff(double ci, double si, double cj, double sj) { return __builtin_fma(ci, cj, si * sj); }
constexpr double kFarMinGap = 0.2;
// Every far pair has r^2 = dx^2 + dy^2 > 0.0081: cosh(2 pi dx) - cos(2 pi dy)
// = 2 sinh^2(pi dx) + 2 sin^2(pi dy) <= 2 sinh^2(pi r) < 0.17 for r <= 0.09.
constexpr double kFarMinR2 = 0.0081;

// The nine far images of xy_fast (see below) as sg(dx) F(|dx|) without the
// sign and the factor 2 pi^2, from t = e^{2 pi |dx|} and ti = 1 / t: the
// two within 1.5 cells as the fraction P / Q, the seven beyond as the sum S.
struct FarParts {
    double P, Q, S;
};
SDLIBM_FN FarParts far_image_parts(double dx, double cy, double t, double ti)
{
    // Far images in terms of ad = |dx| and the sign sg of dx (every image
    // of the reference's sum is odd in dx, so the sum is sg * F(ad)):
    //   dx -+ k (k = 1..3) -> sg * ((ad + k) fA_k + (ad - k) fB_k)
    //   -s dx f(dx + 5s)  -> sg * ad (ad - 5) fB_5        (s = -sg)
    //   (1 + s dx) f(dx - 4s) + f(dx + 4s) -> sg * ((1 - ad)(ad + 4) fA_4 + (ad - 4) fB_4)
    // with fX = (X cy - 1) / (X - cy)^2, A_k = cosh(2 pi (ad + k)) =
    // (t e^{2 pi k} + e^{-2 pi k} / t) / 2 and B_k = cosh(2 pi (k - ad)) =
    // (e^{2 pi k} / t + t e^{-2 pi k}) / 2.
    const double ad = sdlibm::fabsd(dx);
    // |u| = ad + 1 and 1 - ad: the two fractions summed as one, P / Q with Q
    // the product of their denominators
    const double hp = 0.5 * sdlibm_exp_2pik[1], hm = 0.5 * sdlibm_exp_m2pik[1];
    const double A = __builtin_fma(t, hp, ti * hm);
    const double B = __builtin_fma(ti, hp, t * hm);
    const double dA = A - cy, dB = B - cy;
    // (ad + 1) X and (ad - 1) Y as single roundings
    const double X = __builtin_fma(A, cy, -1.0), Y = __builtin_fma(B, cy, -1.0);
    const double nA = __builtin_fma(ad, X, X);
    const double nB = __builtin_fma(ad, Y, -Y);
    const double QA = dA * dA, QB = dB * dB;
    const double P = __builtin_fma(nB, QA, nA * QB);
    const double Q = QA * QB;
    // |u| >= 1.5 (k = 2, 3 both sides, the weighted 4th pair, the 5th): the
    // Fourier series in y of one image (the y-derivative of the Poisson
    // kernel), (c cy - 1) / (c - cy)^2 = 2 sum_n n q^n T_n(cy) with
    // q = e^{-2 pi |u|} and T_n the Chebyshev polynomials (T_n(cy) =
    // cos(2 pi n dy)).  q = E_k ti for the images ad + k and E_k t for
    // k - ad (E_k = e^{-2 pi k}), so the weighted sum over the images is
    //   2 sum_n n T_n(cy) (ti^n Pp_n(ad) + t^n Pm_n(ad)),
    // Pp_n = sum_k w+_k(ad) E_k^n, Pm_n = sum_k w-_k(ad) E_k^n, with the
    // weights w of the image list above.  q <= e^{-3 pi} (k = 2),
    // e^{-5 pi} (k = 3), e^{-7 pi} (k = 4, 5): n <= 4 for k = 2, n <= 2 for
    // k = 3 and n = 1 beyond; the omitted terms are below 3e-17 of the
    // field after the factor 2 pi^2.
    constexpr double E2 = sdlibm_exp_m2pik[2], E3 = sdlibm_exp_m2pik[3], E4 = sdlibm_exp_m2pik[4];
    constexpr double E5 = sdlibm_exp_m2pik[kImages + 1];
    // n = 1: Pp = (ad + 2) E2 + (ad + 3) E3 + (1 - ad)(ad + 4) E4,
    //        Pm = (ad - 2) E2 + (ad - 3) E3 + (ad - 4) E4 + ad (ad - 5) E5 (times 2)
    constexpr double p10 = 2.0 * (2.0 * E2 + 3.0 * E3 + 4.0 * E4), p11 = 2.0 * (E2 + E3 - 3.0 * E4), p12 = -2.0 * E4;
    constexpr double p20 = -p10, p21 = 2.0 * (E2 + E3 + E4 - 5.0 * E5), p22 = 2.0 * E5;
    // n = 2: (ad +- 2) E2^2 + (ad +- 3) E3^2 (times 4); n = 3, 4: (ad +- 2) E2^n (times 2n)
    constexpr double s2 = 4.0 * (E2 * E2 + E3 * E3), c2k = 4.0 * (2.0 * E2 * E2 + 3.0 * E3 * E3);
    constexpr double C3 = 6.0 * E2 * E2 * E2, C4 = 8.0 * E2 * E2 * E2 * E2;
    const double ti2 = ti * ti, t2 = t * t;
    const double ti3 = ti2 * ti, t3 = t2 * t;
    const double ti4 = ti2 * ti2, t4 = t2 * t2;
    const double p1 = __builtin_fma(__builtin_fma(p12, ad, p11), ad, p10);
    const double p2 = __builtin_fma(__builtin_fma(p22, ad, p21), ad, p20);
    const double first = __builtin_fma(ti, p1, t * p2);
    const double second = __builtin_fma(ti2, __builtin_fma(s2, ad, c2k), t2 * __builtin_fma(s2, ad, -c2k));
    const double ap2 = ad + 2.0, am2 = ad - 2.0;
    const double third = __builtin_fma(ti3, ap2, t3 * am2);
    const double fourth = __builtin_fma(ti4, ap2, t4 * am2);
    // sum_{n=1..4} a_n T_n(cy) by Clenshaw's recurrence (a_1 = first,
    // a_2 = second, a_3 = C3 third, a_4 = C4 fourth)
    const double cy2 = 2.0 * cy;
    const double b4 = C4 * fourth;
    const double b3 = __builtin_fma(cy2, b4, C3 * third);
    const double b2 = __builtin_fma(cy2, b3, second - b4);
    const double b1 = __builtin_fma(cy2, b2, first - b3);
    const double S = __builtin_fma(cy, b1, -b2);
    return FarParts{P, Q, S};
}

SDLIBM_FN double far_images(double dx, double cy, double t, double ti)
{
    const FarParts f = far_image_parts(dx, cy, t, ti);
    return f.P / f.Q + f.S;
}

// Production form of xy() (see header comment) with cos(2 pi dy) supplied
// by the caller; xy_fast() below computes it as the reference does.
SDLIBM_FN double xy_fast_cy(double dx, double dy, double cy, const uint64_t *__restrict__ e2tab = sdlibm_exp2_tab)
{
    const double a0 = dx * 2.0 * kPi;
    const double ax = sdlibm::fabsd(a0);
    // Near image: the reference's cosh, evaluated once.  On the exp branch
    // 0.5 / t is taken as 0.5 * (1 / t): scaling by 2^-1 commutes with
    // rounding while 0.5 / t stays normal (t < 2^1021, |a0| < 707; a wrapped
    // dx gives t <= e^pi), so the value is the same and one division serves
    // both uses.
    double c0, t, ti;
    if (sdlibm::hiword(ax) < 0x3fd62e43u) {
        const double m = sdlibm::expm1(ax);
        const double w = 1.0 + m;
        c0 = 1.0 + (m * m) / (w + w);
        t = w;
        ti = 1.0 / t;
    } else {
        t = sdlibm::exp(ax, e2tab);
        ti = 1.0 / t;
        c0 = 0.5 * t + 0.5 * ti;
    }
    const double near = (dx * dx + dy * dy > 1e-6) ? f_given_cosh(dx, c0, cy) : f_taylor(dx, dy);
    // the far images are sg(dx) F(|dx|): the sign rides on the constant
    // (negating one factor of a product negates the rounded product exactly)
    return near + far_images(dx, cy, t, ti) * __builtin_copysign(2.0 * kPiSq, dx);
}

// The production field of a far pair (cosh(2 pi dx) - cos(2 pi dy) >=
// kFarMinGap) in the pair kernel: cos(2 pi dy) and t = e^{2 pi |dx|},
// ti = e^{-2 pi |dx|} supplied by the caller (from per-dislocation factors,
// within a few ulps), cosh(2 pi dx) = (t + ti) / 2.  The near image
// dx (c cy - 1) / (c - cy)^2 and the two inner far images P / Q share one
// division (c - cy >= 0.2 here, so the near image needs no care against
// cancellation).
SDLIBM_FN double xy_far(double dx, double cy, double t, double ti)
{
    // 0.5 t and 0.5 ti are exact, so the fma rounds their sum once, as the
    // sum of the two products does
    const double c0 = __builtin_fma(0.5, t, 0.5 * ti);
    const double d = c0 - cy;
    // every term is odd in dx: evaluated at |dx|, the sign applied with the
    // constant factor (exact: each step's rounding is symmetric in sign)
    const double A = sdlibm::fabsd(dx) * __builtin_fma(c0, cy, -1.0), B = d * d;
    const FarParts f = far_image_parts(dx, cy, t, ti);
    return (sdlibm::div_newton(__builtin_fma(A, f.Q, f.P * B), B * f.Q) + f.S) *
           __builtin_copysign(2.0 * kPiSq, dx);
}

// Per-dislocation factors of the far pairs: the wrapped x w = x - rint(x)
// (exact) and e^{+-2 pi w}.
SDLIBM_FN void xexp(double x, double &w, double &ep, double &em)
{
    w = x - __builtin_rint(x);
    const double a = kPi * 2.0 * w;
    ep = sdlibm::exp(a);
    em = sdlibm::exp(-a);
}

// Circular distance |z - rint(z)| (exact) of a difference of wrapped x.
SDLIBM_FN double circ(double z) { return sdlibm::fabsd(z - __builtin_rint(z)); }

// Bounds of circ(t - s) over t in [tmin, tmax] and s in [smin, smax], the
// intervals widened by 1e-12 for the rounding of how the caller formed them:
// every difference lies in [lo, hi]; circ vanishes at integers, peaks at
// half-integers and is monotone between them, so its extremes over [lo, hi]
// are 0 / 1/2 when the interval holds such a point and otherwise at its ends.
SDLIBM_FN void circ_range(double tmin, double tmax, double smin, double smax, double &mind, double &maxd)
{
    const double lo = tmin - smax - 1e-12, hi = tmax - smin + 1e-12;
    const double clo = circ(lo), chi = circ(hi);
    const bool has_int = __builtin_floor(hi) >= __builtin_ceil(lo);
    const bool has_half = __builtin_floor(hi - 0.5) >= __builtin_ceil(lo - 0.5);
    mind = has_int ? 0.0 : fmin(clo, chi);
    maxd = has_half ? 0.5 : fmax(clo, chi);
}

// How the pairs of a group whose circular x and y distances lie in
// [xmin, xmax] and [ymin, ymax] stand against kFarMinGap: 1 all far, 0 all
// near, 2 either.  With a = pi gx, c = pi gy the gap is 2 sinh^2 a +
// 2 sin^2 c, and a <= sinh a, c (1 - c^2/6) <= sin c <= c (c <= pi/2),
// sinh a <= a (1 + a^2/5) (a <= 1).  The bounds clear kFarMinGap with a
// relative margin of 1e-9, far above their rounding and the far form's.
SDLIBM_FN int gap_kind(double xmin, double xmax, double ymin, double ymax)
{
    const double a0 = kPi * xmin, c0 = kPi * ymin;
    const double a1 = kPi * xmax, c1 = kPi * ymax;
    const double s0 = c0 * (1.0 - c0 * c0 * (1.0 / 6.0));
    const double lower = 2.0 * (a0 * a0 + s0 * s0);
    const double h1 = a1 * (1.0 + a1 * a1 * 0.2);
    const double upper = 2.0 * (h1 * h1 + c1 * c1);
    if (lower >= kFarMinGap * (1.0 + 1e-9)) return 1;
    return (a1 <= 1.0 && upper <= kFarMinGap * (1.0 - 1e-9)) ? 0 : 2;
}

// dx and t = e^{2 pi |dx|}, ti = e^{-2 pi |dx|} of a pair from the two
// dislocations' wrapped x and exp factors: dx = s - rint(s) with s = wi - wj,
// and e^{2 pi dx} = e^{2 pi wi} e^{-2 pi wj} e^{-2 pi k} (k = rint(s) in
// {-1, 0, 1}), each exponential within a few ulps.  u = e^{2 pi s} and
// v = e^{-2 pi s}: hi = max(u, v) = e^{2 pi |s|}, lo = min(u, v).  |dx| is
// min(|s|, 1 - |s|) for |s| < 1, so t = min(hi, lo e^{2 pi}) and
// ti = max(lo, hi e^{-2 pi}) -- without selecting on k, on the sign of dx
// or on whether s wrapped (where |s| is near 1/2 the two candidates agree to
// a few ulps).  Swapping the two dislocations negates s, k and dx exactly
// and exchanges u and v, so t and ti are the same and dx flips sign.
struct FarPrelude {
    double dx, t, ti;
};
SDLIBM_FN FarPrelude far_prelude(double wi, double wj, double epi, double emi, double epj, double emj)
{
    const double s = wi - wj;
    const double u = epi * emj, v = emi * epj;
    const double hi = fmax(u, v), lo = fmin(u, v);
    FarPrelude f;
    f.dx = s - __builtin_rint(s);
    f.t = fmin(hi, lo * sdlibm_exp_2pik[1]);
    f.ti = fmax(lo, hi * sdlibm_exp_m2pik[1]);
    return f;
}

// cosh(2 pi dx) - cos(2 pi dy) as the far form computes it (symmetric in the
// two dislocations bit for bit): the pair is far when it is >= kFarMinGap.
SDLIBM_FN double far_gap(const FarPrelude &f, double cy) { return __builtin_fma(0.5, f.t, 0.5 * f.ti) - cy; }

// The production field of a far pair from the two dislocations' wrapped x,
// exp factors and cos(2 pi dy); odd bit for bit under swapping them.
SDLIBM_FN double far_pair(double wi, double wj, double epi, double emi, double epj, double emj, double cy)
{
    const FarPrelude f = far_prelude(wi, wj, epi, emi, epj, emj);
    return xy_far(f.dx, cy, f.t, f.ti);
}

// The minimum-image term of the reference's sum (AnalyticField.cpp:24-39)
// exactly as xy_fast_cy forms it, for a pair whose cosh(2 pi dx) -
// cos(2 pi dy) is below kFarMinGap: cos(2 pi dy) as glibc forms it (such a
// pair has cos(2 pi dy) > 0.8 - 1e-15, so |2 pi dy| < 0.6436 lies in glibc's
// first range, cos_r1) and cosh(2 pi dx) likewise (expm1 below 0.5 ln 2,
// exp above), then f or its Taylor branch.  Each cosh branch has one
// division (m^2 / 2w, or 1 / t for 0.5 t + 0.5 / t); it is taken once after
// the branches, so lanes that diverge over them share it.  Odd in dx bit
// for bit.
SDLIBM_FN double near_image(double dx, double dy, const double *__restrict__ sctab,
                            const uint64_t *__restrict__ e2tab)
{
    const double cy = sdlibm::cos_r1(kPi * 2.0 * dy, sctab);
    const double ax = sdlibm::fabsd(dx * 2.0 * kPi);
    const bool small = sdlibm::hiword(ax) < 0x3fd62e43u;
    double num, den, t = 0.0;
    if (small) {
        const double m = sdlibm::expm1(ax);
        const double w = 1.0 + m;
        num = m * m;
        den = w + w;
    } else {
        t = sdlibm::exp(ax, e2tab);
        num = 1.0;
        den = t;
    }
    const double q = num / den;
    const double c0 = small ? 1.0 + q : 0.5 * t + 0.5 * q;
    return (dx * dx + dy * dy > 1e-6) ? f_given_cosh(dx, c0, cy) : f_taylor(dx, dy);
}

// The production field of a pair of a mixed tile (some pairs far, some
// near) from the two dislocations' far-pair factors (prelude f, cy by angle
// addition) and their minimum-image differences (dxw, dyw; read only for a
// near pair).  A far pair (gap cosh(2 pi dx) - cos(2 pi dy) >= kFarMinGap,
// as far_gap forms it) gets xy_far's value bit for bit.  A near pair gets
// the minimum image exactly as the reference forms it (near_image, the one
// image whose cosh - cos can cancel) plus the nine far images from the same
// factors as a far pair's (P / Q + S of far_image_parts; every far image has
// |u| >= 1/2, so cosh(2 pi u) - cos >= 10.5 and few-ulp factors move it by
// a few ulps of itself).  Both kinds share far_image_parts and one
// Newton-refined division, so a wavefront with near and far lanes runs the
// far images once and only the near image diverges.  far is set for a far
// pair; r2 = dxw^2 + dyw^2 for a near pair (the tolerance candidates), 1
// otherwise.  Odd under swapping the dislocations bit for bit.
SDLIBM_FN double mixed_pair(const FarPrelude &f, double cy, double dxw, double dyw, bool &far, double &r2,
                            const double *__restrict__ sctab, const uint64_t *__restrict__ e2tab)
{
    const double c0 = __builtin_fma(0.5, f.t, 0.5 * f.ti);
    const double d = c0 - cy;
    const double B = d * d;
    const FarParts p = far_image_parts(f.dx, cy, f.t, f.ti);
    double num = __builtin_fma(sdlibm::fabsd(f.dx) * __builtin_fma(c0, cy, -1.0), p.Q, p.P * B), den = B * p.Q;
    double near = 0.0;
    far = d >= kFarMinGap;
    r2 = 1.0;
    if (!far) {
        near = near_image(dxw, dyw, sctab, e2tab);
        r2 = dxw * dxw + dyw * dyw;
        num = p.P;
        den = p.Q;
    }
    return (sdlibm::div_newton(num, den) + p.S) * __builtin_copysign(2.0 * kPiSq, f.dx) + near;
}

// Production form of xy() (see header comment).
SDLIBM_FN double xy_fast(double dx, double dy, const double *__restrict__ sctab = sdlibm_sincos_tab,
                         const uint64_t *__restrict__ e2tab = sdlibm_exp2_tab)
{
    return xy_fast_cy(dx, dy, sdlibm::cos(kPi * 2.0 * dy, sctab), e2tab);
}

// Production form of xy_diff_x() (AnalyticField.cpp:162-181).  tau_x is
// even in (dx, dy), so with ad = |dx| the sum is
//   near + sum_{k=1,2,3} [Fd(ad + k) + Fd(k - ad)] + (1 - ad) Fd(ad + 4)
//        - F(ad + 4) + Fd(4 - ad) + ad Fd(5 - ad) - F(5 - ad)
// with F(u) = u (c cy - 1) / d^2, Fd(u) = (c cy - 1) / d^2 +
// u sh (2 pi cy / d^2 - 4 pi (c cy - 1) / d^3), c = cosh(2 pi u),
// sh = sinh(2 pi u), d = c - cy (times 2 pi^2).  The near image is the
// reference's f_dx with cosh and sinh formed exactly as glibc forms them
// (one expm1, and one exp where cosh needs it).  All images take cosh and
// sinh from t = e^{2 pi ad}; the
// four within 2.5 cells are summed as one fraction (one division), the five
// beyond by their expansions in a = 1 / cosh (a <= 3e-7, remainder below
// 1e-17 of the term):
//   Fd = a (cy + a c2) - 2 pi u a (cy + 2 a c2),  F = u a (cy + a c2),
// c2 = 2 cy^2 - 1, a = 2 e^{-2 pi u}.
SDLIBM_FN double xy_diff_x_fast(double dx, double dy)
{
    const double cy = sdlibm::cos(kPi * 2.0 * dy);
    const double a0 = kPi * 2.0 * dx;
    const double ax = sdlibm::fabsd(a0);
    // glibc's sinh from expm1(|x|) (|x| < 22), its cosh from expm1 below
    // 0.5 ln 2 and from exp above, both exactly as glibc forms them
    const uint32_t hx = sdlibm::hiword(ax);
    const double m = sdlibm::expm1(ax);
    const double sh0 = hx < 0x3e300000u ? ax
                       : hx < 0x3ff00000u ? 0.5 * (2.0 * m - m * m / (m + 1.0))
                                          : 0.5 * (m + m / (m + 1.0));
    double c0, t;
    if (hx < 0x3fd62e43u) {
        const double w = 1.0 + m;
        c0 = 1.0 + (m * m) / (w + w);
        t = w;
    } else {
        t = sdlibm::exp(ax);
        c0 = 0.5 * t + 0.5 * (1.0 / t);
    }
    const double ti = 1.0 / t;
    double near;
    if (dx * dx + dy * dy > 1e-6) {
        const double sh = a0 < 0 ? -sh0 : sh0;
        const double d = c0 - cy;
        const double d2 = d * d;
        const double d3 = sdlibm::powi(d, 3);
        near = ((c0 * cy - 1.0) / d2 + dx * (sh * cy * 2.0 * kPi / d2 - (c0 * cy - 1.0) / d3 * 4.0 * kPi * sh)) *
               kPi * kPi * 2.0;
    } else {
        near = f_dx_taylor(dx, dy);
    }
    const double ad = sdlibm::fabsd(dx);
    // u = ad + k and k - ad for k = 1, 2: Fd = (P d + u sh (2 pi cy d - 4 pi P)) / d^3
    double N = 0.0, D = 1.0;
    for (int k = 1; k <= 2; ++k) {
        const double hp = 0.5 * sdlibm_exp_2pik[k], hm = 0.5 * sdlibm_exp_m2pik[k];
        for (int side = 0; side < 2; ++side) {
            const double c = side ? __builtin_fma(ti, hp, t * hm) : __builtin_fma(t, hp, ti * hm);
            const double sh = side ? __builtin_fma(ti, hp, -(t * hm)) : __builtin_fma(t, hp, -(ti * hm));
            const double u = side ? (double)k - ad : ad + (double)k;
            const double d = c - cy;
            const double P = __builtin_fma(c, cy, -1.0);
            const double num = __builtin_fma(P, d, (u * sh) * ((2.0 * kPi) * cy * d - (4.0 * kPi) * P));
            const double d3 = d * d * d;
            N = __builtin_fma(num, D, N * d3);
            D *= d3;
        }
    }
    double far = N / D;
    // u >= 2.5: expansions in a = 2 e^{-2 pi u}
    const double c2 = __builtin_fma(2.0 * cy, cy, -1.0);
    const double twopi = 2.0 * kPi;
    const double e3 = 2.0 * sdlibm_exp_m2pik[3], e4 = 2.0 * sdlibm_exp_m2pik[4];
    const double e5 = 2.0 * sdlibm_exp_m2pik[kImages + 1];
    const double us[5] = {ad + 3.0, 3.0 - ad, ad + 4.0, 4.0 - ad, 5.0 - ad};
    const double as[5] = {ti * e3, t * e3, ti * e4, t * e4, t * e5};
    const double wd[5] = {1.0, 1.0, 1.0 - ad, 1.0, ad};     // weights of Fd
    const double wf[5] = {0.0, 0.0, -1.0, 0.0, -1.0};       // weights of F
    double S = 0.0;
    for (int q = 0; q < 5; ++q) {
        const double a = as[q], u = us[q];
        const double g1 = a * __builtin_fma(a, c2, cy);
        const double g2 = a * __builtin_fma(2.0 * a, c2, cy);
        const double fd = __builtin_fma(-(twopi * u), g2, g1);
        S = __builtin_fma(wd[q], fd, S);
        S = __builtin_fma(wf[q] * u, g1, S);
    }
    far += S;
    return near + far * (2.0 * kPiSq);
}

// simulation.cpp:288-322: x-derivative of the point-defect force on a
// dislocation at (dx, dy) from a defect (the Maple expression, grouped as
// three terms that keep its evaluation order), without b and the taper.
SDLIBM_FN double pd_force_dx(double dx, double dy, double A, double K)
{
    const double Pm2 = 0x1.9f02f6222c720p-4;   // pow(M_PI, -2) of the double M_PI, rounded
    const double Pm3 = 0x1.08345eb45d77fp-5;   // pow(M_PI, -3) of the double M_PI, rounded
    const double cx = sdlibm::cos(0.2e1 * kPi * dx), cyy = sdlibm::cos(0.2e1 * kPi * dy);
    const double sx = sdlibm::sin(0.2e1 * kPi * dx), sy = sdlibm::sin(0.2e1 * kPi * dy);
    const double R = (0.1e1 - cx) * Pm2 / 0.2e1 + (0.1e1 - cyy) * Pm2 / 0.2e1;
    const double E = sdlibm::pow_e(-K * R);
    const double logMe = 1.0;                   // log(M_E) rounds to 1
    const double t1 = -A * cx / kPi * sy * ((0.1e1 - E) / R - K * E) / R;
    const double t2 = -A * sx * Pm2 * sy *
                      (E * K * sx / kPi * logMe / R - (0.1e1 - E) * sdlibm::pow_m2(R) * sx / kPi +
                       K * K * E * sx / kPi * logMe) /
                      R / 0.2e1;
    const double t3 = A * (sx * sx) * Pm3 * sy * ((0.1e1 - E) / R - K * E) * sdlibm::pow_m2(R) / 0.2e1;
    return t1 + t2 + t3;
}

}  // namespace sdfield
