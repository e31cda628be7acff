// Host build of the device math headers (libm_compat.h, field.h) for unit
// tests that compare the library's elementary functions and per-pair field
// values with the host C library and the oracle on millions of arguments.
// Test infrastructure: nothing in the shipped library uses this file.
#include "../../sdddst_amd/csrc/field.h"

extern "C" {
double mc_cos(double x) { return sdlibm::cos(x); }
double mc_sin(double x) { return sdlibm::sin(x); }
double mc_exp(double x) { return sdlibm::exp(x); }
double mc_expm1(double x) { return sdlibm::expm1(x); }
double mc_cosh(double x) { return sdlibm::cosh(x); }
double mc_sinh(double x) { return sdlibm::sinh(x); }
double mc_powi(double x, int n) { return sdlibm::powi(x, n); }
double mc_normalize(double d) { return sdfield::normalize(d); }
double mc_xy(double dx, double dy) { return sdfield::xy(dx, dy); }
double mc_xy_fast(double dx, double dy) { return sdfield::xy_fast(dx, dy); }
double mc_xy_diff_x(double dx, double dy) { return sdfield::xy_diff_x(dx, dy); }
double mc_xy_diff_x2(double dx, double dy) { return sdfield::xy_diff_x2(dx, dy); }
double mc_cos_2pi(double p) { return sdfield::cos_2pi(p); }
double mc_pd_force_dx(double dx, double dy, double A, double K) { return sdfield::pd_force_dx(dx, dy, A, K); }

// Batched forms: fn = 0 cos, 1 sin, 2 exp, 3 expm1, 4 cosh, 5 sinh, 6 normalize
void mc_batch1(int fn, long n, const double *x, double *out)
{
    for (long i = 0; i < n; ++i) {
        double v = x[i], r = 0;
        switch (fn) {
        case 0: r = sdlibm::cos(v); break;
        case 1: r = sdlibm::sin(v); break;
        case 2: r = sdlibm::exp(v); break;
        case 3: r = sdlibm::expm1(v); break;
        case 4: r = sdlibm::cosh(v); break;
        case 5: r = sdlibm::sinh(v); break;
        case 6: r = sdfield::normalize(v); break;
        case 9: r = sdlibm::pow_e(v); break;
        case 10: r = sdlibm::pow_m2(v); break;
        case 11: r = sdlibm::cos_r1(v); break;
        }
        out[i] = r;
    }
}

// The pair kernel's far-pair term (its pairs with far_gap >= kFarMinGap):
// cos by angle addition and the exponentials from per-dislocation factors.
void mc_batch_pair_far(long n, const double *xi, const double *yi, const double *xj, const double *yj,
                       double *out)
{
    for (long k = 0; k < n; ++k) {
        double ci, si, cj, sj, wi, epi, emi, wj, epj, emj;
        sdfield::ytrig(yi[k], ci, si);
        sdfield::ytrig(yj[k], cj, sj);
        sdfield::xexp(xi[k], wi, epi, emi);
        sdfield::xexp(xj[k], wj, epj, emj);
        out[k] = sdfield::far_pair(wi, wj, epi, emi, epj, emj, sdfield::cos_diff(ci, si, cj, sj));
    }
}

// The pair kernel's form for the pairs of its mixed tiles (far pairs: the
// far form; near pairs: the reference's minimum image plus the far images
// from the per-dislocation factors), with the far decision in far_out.
void mc_batch_pair_mixed(long n, const double *xi, const double *yi, const double *xj, const double *yj,
                         double *out, int *far_out)
{
    for (long k = 0; k < n; ++k) {
        double ci, si, cj, sj, wi, epi, emi, wj, epj, emj;
        sdfield::ytrig(yi[k], ci, si);
        sdfield::ytrig(yj[k], cj, sj);
        sdfield::xexp(xi[k], wi, epi, emi);
        sdfield::xexp(xj[k], wj, epj, emj);
        const sdfield::FarPrelude f = sdfield::far_prelude(wi, wj, epi, emi, epj, emj);
        bool far;
        double r2;
        out[k] = sdfield::mixed_pair(f, sdfield::cos_diff(ci, si, cj, sj), sdfield::wrap_fast(xi[k] - xj[k]),
                                     sdfield::wrap_fast(yi[k] - yj[k]), far, r2, sdlibm_sincos_tab, sdlibm_exp2_tab);
        far_out[k] = far ? 1 : 0;
    }
}

// The far form's gap cosh(2 pi dx) - cos(2 pi dy) of each pair, from which
// the pair kernel decides far (>= kFarMinGap) or near.
void mc_batch_pair_gap(long n, const double *xi, const double *yi, const double *xj, const double *yj,
                       double *out)
{
    for (long k = 0; k < n; ++k) {
        double ci, si, cj, sj, wi, epi, emi, wj, epj, emj;
        sdfield::ytrig(yi[k], ci, si);
        sdfield::ytrig(yj[k], cj, sj);
        sdfield::xexp(xi[k], wi, epi, emi);
        sdfield::xexp(xj[k], wj, epj, emj);
        const sdfield::FarPrelude f = sdfield::far_prelude(wi, wj, epi, emi, epj, emj);
        out[k] = sdfield::far_gap(f, sdfield::cos_diff(ci, si, cj, sj));
    }
}

// The pair kernel's tile classification for 64 targets and 64 sources
// (x, y as stored; NaN for slots outside the system, never the first): the
// wrapped coordinates shifted by whole periods to within 1/2 of the first
// slot's, their extremes, the circular-distance bounds and gap_kind -- the
// wave reductions of the kernel written as host loops.
static void mc_extent(const double *v, double &vmin, double &vmax)
{
    const double r = v[0] - __builtin_rint(v[0]);
    vmin = vmax = __builtin_nan("");
    for (int l = 0; l < 64; ++l) {
        const double w = v[l] - __builtin_rint(v[l]);
        const double u = w - __builtin_rint(w - r);
        vmin = fmin(vmin, u);
        vmax = fmax(vmax, u);
    }
}
int mc_tile_kind(const double *tx, const double *ty, const double *sx, const double *sy)
{
    double tmin, tmax, smin, smax, xmin, xmax, ymin, ymax;
    mc_extent(tx, tmin, tmax);
    mc_extent(sx, smin, smax);
    sdfield::circ_range(tmin, tmax, smin, smax, xmin, xmax);
    mc_extent(ty, tmin, tmax);
    mc_extent(sy, smin, smax);
    sdfield::circ_range(tmin, tmax, smin, smax, ymin, ymax);
    return sdfield::gap_kind(xmin, xmax, ymin, ymax);
}

// fn = 0 xy, 1 xy_fast, 2 xy_diff_x, 3 xy_diff_x_fast, 4 xy_diff_x2
void mc_batch2(int fn, long n, const double *dx, const double *dy, double *out)
{
    for (long i = 0; i < n; ++i) {
        if (fn == 0) out[i] = sdfield::xy(dx[i], dy[i]);
        else if (fn == 1) out[i] = sdfield::xy_fast(dx[i], dy[i]);
        else if (fn == 2) out[i] = sdfield::xy_diff_x(dx[i], dy[i]);
        else if (fn == 3) out[i] = sdfield::xy_diff_x_fast(dx[i], dy[i]);
        else out[i] = sdfield::xy_diff_x2(dx[i], dy[i]);
    }
}
}
