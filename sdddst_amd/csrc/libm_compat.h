// Elementary functions that reproduce, bit for bit, the results of the host
// C library (glibc 2.35 on x86-64 with FMA, as selected by its ifunc
// dispatch) over the argument ranges the periodic field uses.  The field's
// close-pair terms subtract cos(2*pi*dy) from cosh(2*pi*dx) where both are
// within 1e-5 of 1, so a one-ulp difference in either becomes a 1e-11
// relative difference in the pair term; matching the host library's bits
// removes that source of disagreement entirely.
//
// The methods are the published ones the host library uses:
//   expm1  fdlibm s_expm1.c (Sun Microsystems, 1993), polynomial evaluated
//          in Estrin form, no fused multiply-add;
//   cosh/sinh  fdlibm e_cosh.c / e_sinh.c on top of expm1 and exp;
//   exp    table-driven 2^(k/128) method (Szabolcs Nagy, ARM
//          optimized-routines, 2018), FMA evaluation;
//   sin/cos  IBM Accurate Mathematical Library table method (do_sin /
//          do_cos over sin/cos(i/128)), FMA evaluation, Taylor form for
//          |a| < 0.126, Cody-Waite reduction by pi for |x| > 2.426.
// Tables come from gen_tables.py.  Integer powers use a double-double
// product rounded once (correctly rounded except at exact ties).
//
// Compile with -ffp-contract=off: every fused multiply-add below is written
// out with fma(), and every other product and sum must round on its own.
//
// Under hipcc the functions are device functions; under a host compiler the
// same text builds a host copy for unit tests of the math itself.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SDLIBM_FN __device__ __forceinline__
#define SDLIBM_TABLE static __device__
#else
#include <math.h>
#include <string.h>
#define SDLIBM_FN static inline
#define SDLIBM_TABLE static
#endif

#include "libm_tables.h"

namespace sdlibm {

SDLIBM_FN uint64_t asu64(double d) { return __builtin_bit_cast(uint64_t, d); }
SDLIBM_FN double asf64(uint64_t u) { return __builtin_bit_cast(double, u); }
SDLIBM_FN uint32_t hiword(double d) { return (uint32_t)(asu64(d) >> 32); }
SDLIBM_FN double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
SDLIBM_FN double fabsd(double x) { return __builtin_fabs(x); }
SDLIBM_FN double copysignd(double x, double s) { return __builtin_copysign(x, s); }
// a / b for b > 0 normal and a, a / b finite: the hardware reciprocal
// refined by one Newton step, then one correction of the quotient with the
// exact residual a - b q.  On gfx950 v_rcp_f64 was measured within 2^-24.4
// relative, so the refined reciprocal is within ~2e-15 and the corrected
// quotient within an ulp (it was the correctly rounded quotient for every
// one of 4M random operand pairs over 2^-30..2^30).  For the far-pair form,
// whose error budget is far above an ulp; the host build seeds with the
// division itself.
SDLIBM_FN double div_newton(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
#else
    double r = 1.0 / b;
#endif
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}
// 2^k for -1022 <= k <= 1023, exact.
SDLIBM_FN double pow2i(int k) { return asf64((uint64_t)(1023 + k) << 52); }

// ---------------------------------------------------------------- expm1 --
SDLIBM_FN double expm1(double x)
{
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10,
                 invln2 = 1.44269504088896338700e+00;
    const double Q1 = -3.33333333333331316428e-02, Q2 = 1.58730158725481460165e-03,
                 Q3 = -7.93650757867487942473e-05, Q4 = 4.00821782732936239552e-06,
                 Q5 = -2.01099218183624371326e-07;
    double hi, lo, c = 0.0, t, e, hxs, hfx, r1, y;
    int k;
    uint32_t hx = hiword(x);
    const uint32_t xsb = hx & 0x80000000u;
    hx &= 0x7fffffffu;
    if (hx >= 0x4043687Au) {         // |x| >= 56*ln2
        if (hx >= 0x40862E42u) return x > 0 ? __builtin_inf() : -1.0;
        if (xsb) return -1.0;
    }
    if (hx > 0x3fd62e42u) {          // |x| > 0.5 ln2: reduce
        if (hx < 0x3FF0A2B2u) {      // |x| < 1.5 ln2
            if (!xsb) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
            else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
        } else {
            k = (int)(invln2 * x + (xsb == 0 ? 0.5 : -0.5));
            t = (double)k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (hx < 0x3c900000u) {   // |x| < 2^-54
        return x;
    } else {
        k = 0;
    }
    hfx = 0.5 * x;
    hxs = x * hfx;
    {
        const double R1 = 1.0 + hxs * Q1, h2 = hxs * hxs, R2 = Q2 + hxs * Q3, h4 = h2 * h2, R3 = Q4 + hxs * Q5;
        r1 = R1 + h2 * R2 + h4 * R3;
    }
    t = 3.0 - r1 * hfx;
    e = hxs * ((r1 - t) / (6.0 - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e -= hxs;
    if (k == -1) return 0.5 * (x - e) - 0.5;
    if (k == 1) {
        if (x < -0.25) return -2.0 * (e - (x + 0.5));
        return 1.0 + 2.0 * (x - e);
    }
    if (k <= -2 || k > 56) {
        y = 1.0 - (e - x);
        if (k == 1024) y = y * 2.0 * pow2i(1023);
        else y = y * pow2i(k);
        return y - 1.0;
    }
    if (k < 20) {
        t = 1.0 - pow2i(-k);
        y = t - (e - x);
    } else {
        t = pow2i(-k);
        y = x - (e + t);
        y += 1.0;
    }
    return y * pow2i(k);
}

// ------------------------------------------------------------------ exp --
namespace detail {
// |x| in [512, 1024): the scale's exponent is out of range; shift it and
// rescale (glibc's specialcase; unlike the main path, not fused).
SDLIBM_FN double exp_special(double tmp, uint64_t sbits, uint64_t ki)
{
    if ((ki & 0x80000000u) == 0) {
        sbits -= 1009ull << 52;
        const double scale = asf64(sbits);
        return 0x1p1009 * fmad(scale, tmp, scale);
    }
    sbits += 1022ull << 52;
    const double scale = asf64(sbits);
    double y = scale + scale * tmp;
    if (y < 1.0) {
        double lo = scale - y + scale * tmp;
        const double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) y = 0.0;
    }
    return 0x1p-1022 * y;
}
}  // namespace detail

// exp(x + xt) for a tail |xt| << ulp(x)-scale: the tail joins the reduced
// argument.  exp(x) is exp_tail(x, 0.0), bit for bit.
SDLIBM_FN double exp_tail(double x, double xt, const uint64_t *__restrict__ e2tab = sdlibm_exp2_tab)
{
    const double InvLn2N = 0x1.71547652b82fep0 * 128, NegLn2hiN = -0x1.62e42fefa0000p-8,
                 NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
                 C5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(asu64(x) >> 52) & 0x7ff;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {      // |x| < 2^-54 or |x| >= 512
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;
        if (abstop >= 0x409u) {                    // |x| >= 1024
            if (asu64(x) == asu64(-__builtin_inf())) return 0.0;
            if (abstop >= 0x7ffu) return 1.0 + x;
            return (asu64(x) >> 63) ? 0.0 : __builtin_inf();
        }
        abstop = 0;                                 // large |x|: special case below
    }
    double kd = fmad(InvLn2N, x, Shift);
    const uint64_t ki = asu64(kd);
    kd -= Shift;
    const double r = fmad(kd, NegLn2loN, fmad(kd, NegLn2hiN, x)) + xt;
    const uint64_t idx = 2 * (ki % 128);
    const uint64_t top = ki << 45;
    const double tail = asf64(e2tab[idx]);
    const uint64_t sbits = e2tab[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fmad(r2 * r2, fmad(r, C5, C4), fmad(r2, fmad(r, C3, C2), tail + r));
    if (abstop == 0) return detail::exp_special(tmp, sbits, ki);
    const double scale = asf64(sbits);
    return fmad(scale, tmp, scale);
}

// e2tab / sctab: the tables, or a copy of them (a kernel may stage them in LDS)
SDLIBM_FN double exp(double x, const uint64_t *__restrict__ e2tab = sdlibm_exp2_tab) { return exp_tail(x, 0.0, e2tab); }

// pow(M_E, y) with M_E rounded to double: exp(y * ln(M_E)) with
// ln(M_E) = 1 + c, c = -5.318e-17, i.e. exp of y plus the tail y*c.
SDLIBM_FN double pow_e(double y)
{
    const double c = -0x1.ea8556644e4cdp-55;
    return exp_tail(y, y * c);
}

// pow(x, -2.0) = 1 / x^2 with x^2 kept in double-double, rounded once.
SDLIBM_FN double pow_m2(double x)
{
    const double h = x * x;
    const double l = fmad(x, x, -h);
    const double q = 1.0 / h;
    // 1/(h + l) = q * (1 - (h q - 1) - l q) to first order
    const double e = fmad(h, q, -1.0);
    return fmad(q, -(e + l * q), q);
}

// ----------------------------------------------------------- cosh, sinh --
SDLIBM_FN double cosh(double x)
{
    const uint32_t ix = hiword(x) & 0x7fffffffu;
    const double ax = fabsd(x);
    if (ix < 0x40360000u) {                 // |x| < 22
        if (ix < 0x3fd62e43u) {             // |x| < 0.5 ln2
            if (ix < 0x3c800000u) return 1.0;
            const double t = sdlibm::expm1(ax);
            const double w = 1.0 + t;
            return 1.0 + (t * t) / (w + w);
        }
        const double t = sdlibm::exp(ax);
        return 0.5 * t + 0.5 / t;
    }
    return 0.5 * sdlibm::exp(ax);           // 22 <= |x| < 512
}

SDLIBM_FN double sinh(double x)
{
    const uint32_t jx = hiword(x);
    const uint32_t ix = jx & 0x7fffffffu;
    const double h = (jx & 0x80000000u) ? -0.5 : 0.5;
    const double ax = fabsd(x);
    if (ix < 0x40360000u) {                 // |x| < 22
        if (ix < 0x3e300000u) return x;     // |x| < 2^-28
        const double t = sdlibm::expm1(ax);
        if (ix < 0x3ff00000u) return h * (2.0 * t - t * t / (t + 1.0));
        return h * (t + t / (t + 1.0));
    }
    return h * sdlibm::exp(ax);
}

// -------------------------------------------------------------- sin/cos --
namespace detail {
constexpr double sn3 = -1.66666666666664880952546298448555E-01, sn5 = 8.33333214285722277379541354343671E-03,
                 cs2 = 4.99999999999999999999950396842453E-01, cs4 = -4.16666666666664434524222570944589E-02,
                 cs6 = 1.38888874007937613028114285595617E-03;
constexpr double ts1 = -0.16666666666666666, ts2 = 0.0083333333333323288, ts3 = -0.00019841269834414642,
                 ts4 = 2.755729806860771e-06, ts5 = -2.5022014848318398e-08;
constexpr double big = 52776558133248.0;    // 1.5 * 2^45: rounds |x| to a multiple of 1/128
constexpr double hp0 = 0x1.921fb54442d18p0, hp1 = 0x1.1a62633145c07p-54;    // pi/2 = hp0 + hp1
constexpr double pi_hi = 0x1.921fb54442d18p1, pi_lo = 0x1.1a62633145c07p-53; // pi = pi_hi + pi_lo

SDLIBM_FN int tab_index(double u) { return (int)((uint32_t)asu64(u)) << 2; }

// cos(x + dx) by table lookup at the nearest i/128, |x| < 0.86.
SDLIBM_FN double do_cos(double x, double dx)
{
    if (x < 0) dx = -dx;
    const double u = big + fabsd(x);
    x = fabsd(x) - (u - big) + dx;
    const int k = tab_index(u);
    const double xx = x * x;
    const double s = fmad(x * xx, fmad(xx, sn5, sn3), x);
    const double c = xx * fmad(xx, fmad(xx, cs6, cs4), cs2);
    const double sn = sdlibm_sincos_tab[k], ssn = sdlibm_sincos_tab[k + 1];
    const double cs = sdlibm_sincos_tab[k + 2], ccs = sdlibm_sincos_tab[k + 3];
    const double cor = fmad(-sn, s, fmad(-cs, c, fmad(-s, ssn, ccs)));
    return cs + cor;
}

// sin(x + dx), |x| < 0.86.
SDLIBM_FN double do_sin(double x, double dx)
{
    const double xold = x;
    if (fabsd(x) < 0.126) {
        const double xx = x * x;
        const double P = fmad(fmad(fmad(fmad(ts5, xx, ts4), xx, ts3), xx, ts2), xx, ts1);
        const double t = fmad(fmad(P, x, -0.5 * dx), xx, dx);
        return x + t;
    }
    if (x <= 0) dx = -dx;
    const double u = big + fabsd(x);
    x = fabsd(x) - (u - big);
    const int k = tab_index(u);
    const double xx = x * x;
    const double s = x + fmad(x * xx, fmad(xx, sn5, sn3), dx);
    const double q = fmad(xx, fmad(xx, cs6, cs4), cs2);
    const double c = fmad(x, dx, xx * q);
    const double sn = sdlibm_sincos_tab[k], ssn = sdlibm_sincos_tab[k + 1];
    const double cs = sdlibm_sincos_tab[k + 2], ccs = sdlibm_sincos_tab[k + 3];
    const double cor = fmad(cs, s, fmad(-sn, c, fmad(s, ccs, ssn)));
    return copysignd(sn + cor, xold);
}
}  // namespace detail

// cos(x) for |x| <= 4 (the field evaluates it at 2*pi*d with |d| <= 0.5).
// glibc's three argument ranges -- do_cos(x, 0) below 0.855469,
// do_sin(pi/2 - |x|) below 2.426265 (with its Taylor form below 0.126), and
// -do_cos(|x| - pi) -- are evaluated without divergent branches: one reduced
// argument, one table lookup and the do_cos / do_sin combinations selected
// per lane, each with the operation order of its branch form, so the value
// is the same.  (Random arguments put all three ranges in most wavefronts.)
SDLIBM_FN double cos(double x, const double *__restrict__ sctab = sdlibm_sincos_tab)
{
    using namespace detail;
    const double ax = fabsd(x);
    const uint32_t hk = hiword(ax);
    const bool r1 = hk < 0x3feb6000u;                  // |x| < 0.855469
    const bool r2 = !r1 && hk < 0x400368fdu;           // |x| < 2.426265
    const double y2 = hp0 - ax;
    const double a2 = y2 + hp1;
    const double da2 = (y2 - a2) + hp1;
    const double t3 = ax - pi_hi;                      // exact for |x| in [pi/2, 2pi]
    const double a3 = t3 - pi_lo;
    const double da3 = (t3 - a3) - pi_lo;
    const double a = r1 ? x : (r2 ? a2 : a3);
    const double da0 = r1 ? 0.0 : (r2 ? da2 : da3);
    const double da = (r2 ? a <= 0 : a < 0) ? -da0 : da0;
    const double u = big + fabsd(a);
    const double xr = fabsd(a) - (u - big);
    const int k = tab_index(u);
    const double y = r2 ? xr : xr + da;                // do_sin keeps dx apart
    const double xx = y * y;
    const double P = fmad(xx, sn5, sn3);
    const double s = r2 ? y + fmad(y * xx, P, da) : fmad(y * xx, P, y);
    const double q = xx * fmad(xx, fmad(xx, cs6, cs4), cs2);
    const double c = r2 ? fmad(y, da, q) : q;
    // do_sin combines (sn, ssn) with (cs, ccs), do_cos (cs, ccs) with
    // (-sn, -ssn): pick the table pairs by offset and carry do_cos's minus
    // sign on s (B * s = (-B') * s = B' * (-s) exactly)
    const int ka = k + (r2 ? 0 : 2), kb = k + (r2 ? 2 : 0);
    const double A = sctab[ka], Alo = sctab[ka + 1];
    const double B = sctab[kb], Blo = sctab[kb + 1];
    const double sb = r2 ? s : -s;
    const double res = A + fmad(B, sb, fmad(-A, c, fmad(sb, Blo, Alo)));
    // do_sin's Taylor form for |a| < 0.126 (dx as passed in)
    const double tx = a2 * a2;
    const double tP = fmad(fmad(fmad(fmad(ts5, tx, ts4), tx, ts3), tx, ts2), tx, ts1);
    const double tay = a2 + fmad(fmad(tP, a2, -0.5 * da2), tx, da2);
    const double v2 = fabsd(a2) < 0.126 ? tay : copysignd(res, a2);
    const double v = r1 ? res : (r2 ? v2 : -res);
    return hk < 0x3e400000u ? 1.0 : v;                 // |x| < 2^-27
}

// cos(x) for |x| < 0.855469: glibc's first range alone, do_cos(x, 0) (1
// below 2^-27), the same value as cos() there with none of the other
// ranges' work.  Larger arguments are clamped into the range, so the table
// lookup stays in bounds (their value is then meaningless).
SDLIBM_FN double cos_r1(double x, const double *__restrict__ sctab = sdlibm_sincos_tab)
{
    using namespace detail;
    const double ax = fmin(fabsd(x), 0.8555);   // table entry 110 at most
    const double u = big + ax;
    const double y = ax - (u - big);
    const int k = tab_index(u);
    const double xx = y * y;
    const double s = fmad(y * xx, fmad(xx, sn5, sn3), y);
    const double c = xx * fmad(xx, fmad(xx, cs6, cs4), cs2);
    const double A = sctab[k + 2], Alo = sctab[k + 3], B = sctab[k], Blo = sctab[k + 1];
    const double res = A + fmad(B, -s, fmad(-A, c, fmad(-s, Blo, Alo)));
    return hiword(ax) < 0x3e400000u ? 1.0 : res;
}

// sin(x) for |x| <= 4.
SDLIBM_FN double sin(double x)
{
    using namespace detail;
    const double ax = fabsd(x);
    const uint32_t k = hiword(ax);
    if (k < 0x3e500000u) return x;                     // |x| < 2^-26
    if (k < 0x3feb6000u) return do_sin(x, 0.0);
    if (k < 0x400368fdu) return copysignd(do_cos(hp0 - ax, hp1), x);
    const double t = ax - pi_hi;
    const double a = t - pi_lo;
    const double da = (t - a) - pi_lo;
    return copysignd(-do_sin(a, da), x);
}

// ------------------------------------------------------ integer powers --
// x^n for 2 <= n <= 7 as a double-double product rounded once.
SDLIBM_FN double powi(double x, int n)
{
    double hi = x, lo = 0.0;
    for (int i = 1; i < n; ++i) {
        const double p = hi * x;
        lo = fmad(hi, x, -p) + lo * x;
        hi = p;
    }
    return hi + lo;
}

// IEEE square root (correctly rounded on both sides).
SDLIBM_FN double sqrt(double x) { return __builtin_sqrt(x); }

}  // namespace sdlibm

namespace sdlibm {
// pow(t, 1./3.) for finite t > 0: the cube root refined in double-double,
// times t^(-(1/3 - RN(1/3))) = 1 - ln(t) * 2^-54/3, rounded once.
// (PrecisionHandler::getNewStepSize, precision_handler.cpp:111.)
SDLIBM_FN double pow_one_third(double t)
{
    const double c = ::cbrt(t);
    const double c2 = c * c;
    const double c2l = fmad(c, c, -c2);
    const double c3 = c2 * c;
    const double c3l = fmad(c2, c, -c3) + c2l * c;
    const double r = (t - c3) - c3l;
    const double corr = r / (3.0 * c2);
    const double delta = 0x1p-54 / 3.0;
    const double lo = corr - c * (delta * ::log(t));
    return c + lo;
}
}  // namespace sdlibm
