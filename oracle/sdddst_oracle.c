/*
 * CPU oracle for the SDDDST implicit-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark compare the HIP
 * library against.  Nothing in the shipped library links, loads or calls it.
 *
 * It restates, in plain C and in the reference's operation order, the
 * functions of pgabor/sdddst that sit on the implicit-step path.  Every
 * function cites the reference lines it follows.  Build flags matter for
 * bit-parity: -O2 -ffp-contract=off, no -march (see oracle/Makefile).  With
 * those flags GCC folds pow(x, 2.0) into x*x exactly as it does for an -O2
 * build of the reference; every other libm call (cos, cosh, sinh, exp, pow
 * with exponents 3..7, sqrt) goes to the host glibc, as in the reference.
 *
 * Parity pins (tests/test_oracle.py):
 *   - orc_xy / orc_xy_diff_x / orc_xy_diff_x2 / precision handler:
 *     bit-identical to the reference's own AnalyticField / PrecisionHandler
 *     built into oracle/_ref (when it is present) and to the committed
 *     golden vectors;
 *   - closed forms: y-periodic image sum, odd symmetry, small-r limit,
 *     Taylor-branch continuity, central differences (mpmath);
 *   - speeds: N=2 dipole, sum rule, brute-force long-double sums;
 *   - Jacobian: brute-force predicate set equality, column-sum identity;
 *   - dense LU solve: residual check;
 *   - calculateG / integrate / passes 2-3 / the three stages
 *     (tests/test_oracle_stepping.py): the assembled matrix equals central
 *     differences of the residual at an infinite cutoff (with and without
 *     point defects), the speed weights sum to one, a dipole's linear mode
 *     follows the theta-scheme with theta = (1+d)/2, local error order 3,
 *     second-order convergence to a DOP853 solution, N = 1 closed forms;
 *   - stress protocols: the golden table made from the reference's own
 *     protocol translation units;
 *   - point-defect terms: mpmath derivatives of the speed term;
 *   - the run loop (tests/test_oracle_runloop.py): each limit, the avalanche
 *     counter, the cutoff switch, post-relax and sub-configurations end the
 *     run where the loop's semantics place them on the log of an
 *     unconstrained run, whose logged prefix they reproduce.
 * The dense LU solver below stands in for UMFPACK (absent from this image);
 * its solves are pinned only by residual, "parity unpinned" against UMFPACK.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_ANALYTIC_FIELD_N 4      /* constants.h:25 */
#define ORC_EPS 1e-12               /* constants.h:24 */

/* utility.h:29-40: wrap into [-0.5, 0.5) by repeated +/-1 (loop kept). */
double orc_normalize(double n)
{
    while (n < -0.5) n += 1.0;
    while (n >= 0.5) n -= 1.0;
    return n;
}

/* AnalyticField.cpp:24-39 : one y-periodic wall of images, x offset dx. */
static double fwall(double dx, double cos2piy, double dy)
{
    double cosh2pix;
    if (dx * dx + dy * dy > 1e-6) {
        cosh2pix = cosh(dx * 2.0 * M_PI);
        return dx * (cosh2pix * cos2piy - 1.0) / ((cosh2pix - cos2piy) * (cosh2pix - cos2piy)) * 2. * M_PI * M_PI;
    }
    {
        double dx2pi = dx * 2.0 * M_PI;
        double dy2pi = dy * 2.0 * M_PI;
        double a = pow(dx2pi, 6.0) / 720.0 + pow(dx2pi, 4.0) / 24.0 + pow(dx2pi, 2) * 0.5;
        double bb = -pow(dy2pi, 6.0) / 720.0 + pow(dy2pi, 4) / 24.0 - pow(dy2pi, 2) * 0.5;
        double c = (pow(dy2pi, 6.0) + pow(dx2pi, 6.0)) / 720. + (pow(dx2pi, 4.0) - pow(dy2pi, 4.0)) / 24.0 +
                   (pow(dx2pi, 2.0) + pow(dy2pi, 2.0)) * 0.5;
        return ((a * bb + bb + a) * dx) / (pow(c, 2.0)) * 2.0 * pow(M_PI, 2.0);
    }
}

/* AnalyticField.cpp:41-51 : g<k> = g<k-1> + f(dx-k) + f(dx+k), g<0> = f(dx). */
static double gsum(int images, double dx, double cos2piy, double dy)
{
    double acc = fwall(dx, cos2piy, dy);
    for (int k = 1; k <= images; ++k)
        acc = acc + fwall(dx - (double)k, cos2piy, dy) + fwall(dx + (double)k, cos2piy, dy);
    return acc;
}

/* AnalyticField.cpp:53-93 : d/dx of fwall. */
static double fwall_dx(double dx, double cos2piy, double dy)
{
    double cosh2pix = cosh(M_PI * 2.0 * dx);
    double sinh2pix = sinh(M_PI * 2.0 * dx);

    if (dx * dx + dy * dy > 1e-6) {
        return ((cosh2pix * cos2piy - 1.0) / pow(cosh2pix - cos2piy, 2.0) +
                dx * (sinh2pix * cos2piy * 2.0 * M_PI / pow(cosh2pix - cos2piy, 2.0) -
                      (cosh2pix * cos2piy - 1.0) / pow(cosh2pix - cos2piy, 3.0) * 4.0 * M_PI * sinh2pix)) *
               M_PI * M_PI * 2.0;
    }
    {
        double X = dx * 2.0 * M_PI;
        double Y = dy * 2.0 * M_PI;
        double a = pow(X, 6.0) / 720.0 + pow(X, 4.0) / 24.0 + pow(X, 2) * 0.5;
        double bb = -pow(Y, 6.0) / 720.0 + pow(Y, 4) / 24.0 - pow(Y, 2) * 0.5;
        double c = (pow(Y, 6.0) + pow(X, 6.0)) / 720. + (pow(X, 4.0) - pow(Y, 4.0)) / 24.0 +
                   (pow(X, 2.0) + pow(Y, 2.0)) * 0.5;
        double cs = -(pow(X, 7.0) * pow(Y, 6.0)) / 3628800.0
                    - (pow(X, 5.0) * pow(Y, 6.0)) / 86400.0
                    - (pow(X, 3.0) * pow(Y, 6.0)) / 4320.0
                    - (pow(Y, 6.0) * X) / 720
                    + (pow(X, 7.0) * pow(Y, 4.0)) / 120960.0
                    + (pow(X, 5.0) * pow(Y, 4.0)) / 2880.0
                    + (pow(X, 3.0) * pow(Y, 4.0)) / 144.0
                    + (pow(Y, 4.0) * X) / 24.0
                    - (pow(X, 7.0) * pow(Y, 2.0)) / 1080.0
                    - (pow(X, 5.0) * pow(Y, 2.0)) / 240.0
                    - (pow(X, 3.0) * pow(Y, 2.0)) / 12.0
                    - (pow(Y, 2.0) * X) * 0.5
                    + pow(X, 7.0) / 5040.0
                    + pow(X, 5.0) / 120.0
                    + pow(X, 3.0) / 6.0
                    + X;
        return ((a * bb + bb + a) / pow(c, 2.0) +
                ((cs * dx) / (pow(c, 2.0)) - (a * bb + bb + a) / (pow(c, 3.0)) * dx * sinh(X) * 2.0) * M_PI * 2.0) *
               2.0 * pow(M_PI, 2.0);
    }
}

/* AnalyticField.cpp:110-120 */
static double gsum_dx(int images, double dx, double cos2piy, double dy)
{
    double acc = fwall_dx(dx, cos2piy, dy);
    for (int k = 1; k <= images; ++k)
        acc = acc + fwall_dx(dx - (double)k, cos2piy, dy) + fwall_dx(dx + (double)k, cos2piy, dy);
    return acc;
}

/* AnalyticField.cpp:145-160 : periodic shear stress tau_xy(dx, dy). */
double orc_xy(double dx, double dy)
{
    const double Nf = (double)ORC_ANALYTIC_FIELD_N;
    double cos2piy = cos(M_PI * 2.0 * dy);
    if (dx < 0) {
        return -dx * fwall(dx + (double)(ORC_ANALYTIC_FIELD_N + 1), cos2piy, dy) +
               (1.0 + dx) * fwall(dx - Nf, cos2piy, dy) +
               fwall(dx + Nf, cos2piy, dy) +
               gsum(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
    }
    return dx * fwall(dx - (double)(ORC_ANALYTIC_FIELD_N + 1), cos2piy, dy) +
           (1.0 - dx) * fwall(dx + Nf, cos2piy, dy) +
           fwall(dx - Nf, cos2piy, dy) +
           gsum(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
}

/* AnalyticField.cpp:162-181 : d tau_xy / dx. */
double orc_xy_diff_x(double dx, double dy)
{
    const double Nf = (double)ORC_ANALYTIC_FIELD_N;
    double cos2piy = cos(M_PI * 2.0 * dy);
    if (dx < 0) {
        return -fwall(dx + Nf + 1.0, cos2piy, dy) -
               dx * fwall_dx(dx + Nf + 1.0, cos2piy, dy) +
               fwall_dx(dx - Nf, cos2piy, dy) * (1.0 + dx) +
               fwall(dx - Nf, cos2piy, dy) +
               fwall_dx(dx + Nf, cos2piy, dy) +
               gsum_dx(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
    }
    return fwall(dx - Nf - 1.0, cos2piy, dy) +
           dx * fwall_dx(dx - Nf - 1.0, cos2piy, dy) +
           fwall_dx(dx + Nf, cos2piy, dy) * (1.0 - dx) -
           fwall(dx + Nf, cos2piy, dy) +
           fwall_dx(dx - Nf, cos2piy, dy) +
           gsum_dx(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
}

/* AnalyticField.cpp:95-106 : second x-derivative of fwall (no small-r
 * branch: the reference divides by (cos 2 pi dy - cosh 2 pi dx)^4 for every
 * image). */
static double fwall_dx2(double dx, double cos2piy, double dy)
{
    double ch = cosh(M_PI * 2.0 * dx);
    double sh = sinh(M_PI * 2.0 * dx);
    double c4y = cos(4.0 * M_PI * dy);
    double ch2 = cosh(M_PI * 4.0 * dx);
    double sh2 = sinh(M_PI * 4.0 * dx);
    double num = sh * cos2piy * (pow(cos2piy, 2.0) - 2.0) +
                 2.0 * M_PI * dx * pow(sh, 2.0) * (c4y - 2.0) -
                 M_PI * dx * pow(ch, 3.0) * cos2piy +
                 0.5 * M_PI * dx * ch * cos2piy * (2.0 * ch2 + c4y - 5.0) +
                 pow(ch, 2.0) * (2.0 * M_PI * dx - sh * cos2piy) +
                 sh2;
    return (4.0 * M_PI * num) / pow(cos2piy - ch, 4.0);
}

/* AnalyticField.cpp:122-132 */
static double gsum_dx2(int images, double dx, double cos2piy, double dy)
{
    double acc = fwall_dx2(dx, cos2piy, dy);
    for (int k = 1; k <= images; ++k)
        acc = acc + fwall_dx2(dx - (double)k, cos2piy, dy) + fwall_dx2(dx + (double)k, cos2piy, dy);
    return acc;
}

/* AnalyticField.cpp:183-202 : d^2 tau_xy / dx^2 (the eigenvalue-derivative
 * option of the EV analysis). */
double orc_xy_diff_x2(double dx, double dy)
{
    const double Nf = (double)ORC_ANALYTIC_FIELD_N;
    double cos2piy = cos(M_PI * 2.0 * dy);
    if (dx < 0) {
        return -2.0 * fwall_dx(dx + Nf + 1.0, cos2piy, dy) -
               dx * fwall_dx2(dx + Nf + 1.0, cos2piy, dy) +
               fwall_dx2(dx - Nf, cos2piy, dy) * (1.0 + dx) +
               2.0 * fwall_dx(dx - Nf, cos2piy, dy) +
               fwall_dx2(dx + Nf, cos2piy, dy) +
               gsum_dx2(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
    }
    return 2.0 * fwall_dx(dx - Nf - 1.0, cos2piy, dy) +
           dx * fwall_dx2(dx - Nf - 1.0, cos2piy, dy) +
           fwall_dx2(dx + Nf, cos2piy, dy) * (1.0 - dx) -
           2.0 * fwall_dx(dx + Nf, cos2piy, dy) +
           fwall_dx2(dx - Nf, cos2piy, dy) +
           gsum_dx2(ORC_ANALYTIC_FIELD_N - 1, dx, cos2piy, dy);
}

/* utility.h:42-55 : point-defect helpers. */
static double pX(double x) { return sin(2.0 * M_PI * x) * 0.5 / M_PI; }
static double pX2(double x) { return (1.0 - cos(2.0 * M_PI * x)) * 0.5 / M_PI / M_PI; }

/* ---------------------------------------------------------------------- */
/* PrecisionHandler (precision_handler.cpp:28-119)                         */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n;
    double *tol;       /* toleranceAndError[i].first  */
    double *err;       /* toleranceAndError[i].second */
    double minPrec, minPrecSqr, maxErrorRatioSqr;
    unsigned selectedID;
    long samePlaceMessages; /* count of the stdout message at :80 */
} orc_ph;

/* precision_handler.cpp:72-87 */
static void ph_update_tolerance(orc_ph *p, double distanceSqr, unsigned id)
{
    double tmp = distanceSqr * 0.25 * 1e-2;
    if (tmp < p->minPrecSqr && tmp < p->tol[id]) {
        if (tmp == 0.0) {
            p->tol[id] = ORC_EPS;
            p->samePlaceMessages++;
        } else {
            p->tol[id] = tmp;
        }
    }
}

/* precision_handler.cpp:89-101 */
static void ph_update_error(orc_ph *p, double error, unsigned id)
{
    if (p->err[id] < error) {
        p->err[id] = error;
        double tmp = error * error / p->tol[id];
        if (tmp > p->maxErrorRatioSqr) {
            p->maxErrorRatioSqr = tmp;
            p->selectedID = id;
        }
    }
}

/* precision_handler.cpp:103-119 */
double orc_new_step_size(double maxErrorRatioSqr, double oldStepSize)
{
    if (0.0 == maxErrorRatioSqr) return oldStepSize * 2.0;
    double tmp = 1. / sqrt(maxErrorRatioSqr);
    tmp = pow(tmp, 1. / 3.);
    if (tmp > 2.0) tmp = 2.0;
    return 0.9 * oldStepSize * tmp;
}

/* precision_handler.cpp:62-70 */
static void ph_reset(orc_ph *p)
{
    p->maxErrorRatioSqr = 0;
    for (int i = 0; i < p->n; ++i) {
        p->tol[i] = p->minPrecSqr;
        p->err[i] = 0;
    }
}

/* Stand-alone entry points over caller-owned arrays (used by the tests). */
void orc_ph_update_tolerance(double *tol, double minPrecSqr, double distanceSqr, unsigned id)
{
    orc_ph p;
    memset(&p, 0, sizeof p);
    p.tol = tol;
    p.minPrecSqr = minPrecSqr;
    ph_update_tolerance(&p, distanceSqr, id);
}

/* simulation.cpp:398-405 + precision_handler.cpp:89-101 over one step:
 * err starts at 0 (reset), so each id is visited exactly once. */
double orc_max_error_ratio(int n, const double *xbig, const double *xsmall2, const double *tol, int *selected)
{
    double maxr = 0;
    int sel = 0;
    for (int i = 0; i < n; ++i) {
        double e = fabs(xbig[i] - xsmall2[i]);
        if (0.0 < e) {
            double tmp = e * e / tol[i];
            if (tmp > maxr) {
                maxr = tmp;
                sel = i;
            }
        }
    }
    if (selected) *selected = sel;
    return maxr;
}

/* ---------------------------------------------------------------------- */
/* Simulation::calculateSpeeds (simulation.cpp:86-131)                    */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n, pc;
    const double *x, *y, *b;   /* dislocations (SoA)      */
    const double *px, *py;     /* point defects           */
    double A, KASQR;
} orc_sys;

static void speeds_impl(const orc_sys *s, double sigma, double *res, orc_ph *ph)
{
    const int n = s->n;
    for (int i = 0; i < n; ++i) res[i] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(s->x[i] - s->x[j]);
            double dy = orc_normalize(s->y[i] - s->y[j]);
            double tmp = s->b[i] * s->b[j] * orc_xy(dx, dy);
            double r2 = dx * dx + dy * dy;
            if (ph) {
                ph_update_tolerance(ph, r2, (unsigned)i);
                ph_update_tolerance(ph, r2, (unsigned)j);
            }
            res[i] += tmp;
            res[j] -= tmp;
        }
        for (int j = 0; j < s->pc; ++j) {
            double dx = orc_normalize(s->x[i] - s->px[j]);
            double dy = orc_normalize(s->y[i] - s->py[j]);
            double xSqr = pX2(dx);
            double ySqr = pX2(dy);
            double rSqr = xSqr + ySqr;
            double expXY = exp(-s->KASQR * rSqr);
            res[i] -= 2.0 * s->A * pX(dx) * pX(dy) * ((1. - expXY) / rSqr - s->KASQR * expXY) / rSqr * s->b[i];
            if (ph) ph_update_tolerance(ph, rSqr, (unsigned)i);
        }
        res[i] += s->b[i] * sigma;
    }
}

/* ctypes entry: tol (nullable) is updated in place with minPrecSqr as the
 * precision handler's lower bound (precision_handler.cpp:74). */
void orc_speeds(int n, const double *x, const double *y, const double *b,
                int pc, const double *px, const double *py, double A, double KASQR,
                double sigma, double *v, double *tol, double minPrecSqr)
{
    orc_sys s = {n, pc, x, y, b, px, py, A, KASQR};
    if (tol) {
        orc_ph p;
        memset(&p, 0, sizeof p);
        p.n = n;
        p.tol = tol;
        p.minPrecSqr = minPrecSqr;
        speeds_impl(&s, sigma, v, &p);
    } else {
        speeds_impl(&s, sigma, v, NULL);
    }
}

/* Speeds over a subset of source rows i in [i0, i1) only: the bounded CPU
 * baseline sample (same per-pair work as orc_speeds, no cross-row sums). */
double orc_speeds_rows_sample(int n, const double *x, const double *y, const double *b,
                              int i0, int i1, long *pairs_out)
{
    double acc = 0;
    long pairs = 0;
    for (int i = i0; i < i1; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(x[i] - x[j]);
            double dy = orc_normalize(y[i] - y[j]);
            acc += b[i] * b[j] * orc_xy(dx, dy);
            ++pairs;
        }
    if (pairs_out) *pairs_out = pairs;
    return acc;
}

/* ---------------------------------------------------------------------- */
/* Simulation::calculateJacobian (simulation.cpp:195-384)                   */
/* ---------------------------------------------------------------------- */
typedef struct {
    double cutOff, cutOffSqr, onePerCutOffSqr;
} orc_cut;

/* simulation_data.cpp:208-218 */
void orc_cutoff(int n, double cutOffMultiplier, double *cutOff, double *cutOffSqr, double *onePerCutOffSqr)
{
    double multiplier = cutOffMultiplier;
    if (cutOffMultiplier >= 1. / 12. * sqrt(2.0 * (double)n)) multiplier = 1e20;
    *cutOff = 1. / sqrt((double)n) * multiplier;
    *cutOffSqr = *cutOff * *cutOff;
    *onePerCutOffSqr = 1. / *cutOffSqr;
}

/* simulation.cpp:195-229 : recursive binary search of row j in [si, ei). */
static double get_element(const int *Ai, const double *Ax, int j, int si, int ei)
{
    int len = ei - si;
    if (len > 1) {
        int tmp = len / 2;
        double a;
        if (Ai[si + tmp] > j) {
            a = get_element(Ai, Ax, j, si, si + tmp);
            if (a != 0.0) return a;
        } else {
            a = get_element(Ai, Ax, j, si + tmp, ei);
            if (a != 0.0) return a;
        }
    } else {
        if (Ai[si] == j) return Ax[si];
        return 0;
    }
    return 0;
}

/* Point-defect diagonal term of simulation.cpp:274-324 (value added to tmp). */
static double pd_diag_term(double dx, double dy, double bj, double A, double K, double multiplier)
{
    const double Pm2 = pow(M_PI, -0.2e1);
    double cx = cos(0.2e1 * M_PI * dx), cy = cos(0.2e1 * M_PI * dy);
    double sx = sin(0.2e1 * M_PI * dx), sy = sin(0.2e1 * M_PI * dy);
    double R = (0.1e1 - cx) * Pm2 / 0.2e1 + (0.1e1 - cy) * Pm2 / 0.2e1;
    double E = pow(M_E, -K * R);
    double t1 = -A * cx / M_PI * sy * ((0.1e1 - E) / R - K * E) / R;
    double t2 = -A * sx * Pm2 * sy *
                (E * K * sx / M_PI * log(M_E) / R - (0.1e1 - E) * pow(R, -0.2e1) * sx / M_PI +
                 K * K * E * sx / M_PI * log(M_E)) / R / 0.2e1;
    double t3 = A * pow(sx, 0.2e1) * pow(M_PI, -0.3e1) * sy * ((0.1e1 - E) / R - K * E) * pow(R, -0.2e1) / 0.2e1;
    return bj * (t1 + t2 + t3) * multiplier;
}

/* ctypes entry: the point-defect term of simulation.cpp:288-322 for b, taper. */
double orc_pd_diag_term(double dx, double dy, double bj, double A, double K, double multiplier)
{
    return pd_diag_term(dx, dy, bj, A, K, multiplier);
}

/* Pass 1 of calculateJacobian (simulation.cpp:240-347).  Fills Ap, Ai, Ax
 * (raw values, diag = -pd*h).  Returns nnz, or -1 if cap is exceeded. */
static long jac_pass1(const orc_sys *s, const orc_cut *c, double stepsize, long cap, int *Ap, int *Ai, double *Ax)
{
    const int n = s->n;
    long cnt = 0;
    Ap[0] = 0;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < j; ++i) {
            double v = get_element(Ai, Ax, j, Ap[i], Ap[i + 1]);
            if (v != 0.0) {
                if (cnt >= cap) return -1;
                Ai[cnt] = i;
                Ax[cnt++] = v;
            }
        }
        if (cnt >= cap) return -1;
        Ai[cnt] = j;
        double tmp = 0;
        for (int l = 0; l < s->pc; ++l) {
            double dx = orc_normalize(s->x[j] - s->px[l]);
            double dy = orc_normalize(s->y[j] - s->py[l]);
            if (pow(sqrt(dx * dx + dy * dy) - c->cutOff, 2) < 36.8 * c->cutOffSqr) {
                double multiplier = 1;
                if (dx * dx + dy * dy > c->cutOffSqr)
                    multiplier = exp(-pow(sqrt(dx * dx + dy * dy) - c->cutOff, 2) * c->onePerCutOffSqr);
                tmp -= pd_diag_term(dx, dy, s->b[j], s->A, s->KASQR, multiplier);
            }
        }
        Ax[cnt++] = -tmp * stepsize;
        for (int i = j + 1; i < n; ++i) {
            double dx = orc_normalize(s->x[i] - s->x[j]);
            double dy = orc_normalize(s->y[i] - s->y[j]);
            if (pow(sqrt(dx * dx + dy * dy) - c->cutOff, 2) < 36.8 * c->cutOffSqr) {
                double multiplier = 1;
                if (dx * dx + dy * dy > c->cutOffSqr)
                    multiplier = exp(-pow(sqrt(dx * dx + dy * dy) - c->cutOff, 2) * c->onePerCutOffSqr);
                if (cnt >= cap) return -1;
                Ai[cnt] = i;
                Ax[cnt++] = stepsize * s->b[i] * s->b[j] * orc_xy_diff_x(dx, dy) * multiplier;
            }
        }
        Ap[j + 1] = (int)cnt;
    }
    return cnt;
}

/* Passes 2 and 3 (simulation.cpp:349-383). */
static void jac_pass23(int n, const int *Ap, const int *Ai, double *Ax, double *dVec, int *indexes)
{
    for (int j = 0; j < n; ++j) {
        double subSum = 0;
        for (int i = Ap[j]; i < Ap[j + 1]; ++i) {
            if (Ai[i] == j) indexes[j] = i;
            subSum += Ax[i];
        }
        subSum *= -1.;
        Ax[indexes[j]] = subSum;
        if (subSum > 0) {
            subSum = 1. / subSum;
            subSum += 1.;
            subSum *= subSum;
            dVec[j] = 1. / subSum;
        } else {
            dVec[j] = 0.;
        }
    }
    for (int j = 0; j < n; ++j) {
        for (int i = Ap[j]; i < Ap[j + 1]; ++i) Ax[i] *= (1.0 + dVec[Ai[i]]) * 0.5;
        Ax[indexes[j]] += 1.0;
    }
}

/* ctypes entry: full calculateJacobian.  raw_Ax (nullable) receives the
 * pass-1 values before passes 2-3.  Returns nnz or -1 when cap is small. */
long orc_jacobian(int n, const double *x, const double *y, const double *b,
                  int pc, const double *px, const double *py, double A, double KASQR,
                  double cutOffMultiplier, double stepsize, long cap,
                  int *Ap, int *Ai, double *Ax, double *dVec, int *indexes, double *raw_Ax)
{
    orc_sys s = {n, pc, x, y, b, px, py, A, KASQR};
    orc_cut c;
    orc_cutoff(n, cutOffMultiplier, &c.cutOff, &c.cutOffSqr, &c.onePerCutOffSqr);
    long nnz = jac_pass1(&s, &c, stepsize, cap, Ap, Ai, Ax);
    if (nnz < 0) return nnz;
    if (raw_Ax) memcpy(raw_Ax, Ax, (size_t)nnz * sizeof(double));
    jac_pass23(n, Ap, Ai, Ax, dVec, indexes);
    return nnz;
}

/* ---------------------------------------------------------------------- */
/* Dense LU with partial pivoting: the UMFPACK stand-in of the oracle.      */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n;
    double *lu;   /* row-major n*n */
    int *piv;
} orc_lu;

static int lu_factor_csc(orc_lu *f, int n, const int *Ap, const int *Ai, const double *Ax)
{
    f->n = n;
    f->lu = (double *)calloc((size_t)n * n, sizeof(double));
    f->piv = (int *)malloc((size_t)n * sizeof(int));
    if (!f->lu || !f->piv) return -1;
    for (int j = 0; j < n; ++j)
        for (int k = Ap[j]; k < Ap[j + 1]; ++k) f->lu[(size_t)Ai[k] * n + j] += Ax[k];
    double *a = f->lu;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[(size_t)i * n + k]) > best) {
                best = fabs(a[(size_t)i * n + k]);
                p = i;
            }
        f->piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)p * n + j];
                a[(size_t)p * n + j] = t;
            }
        double d = a[(size_t)k * n + k];
        if (d == 0.0) return -2;
        for (int i = k + 1; i < n; ++i) {
            double l = a[(size_t)i * n + k] / d;
            a[(size_t)i * n + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
        }
    }
    return 0;
}

static void lu_solve(const orc_lu *f, const double *rhs, double *out)
{
    int n = f->n;
    const double *a = f->lu;
    for (int i = 0; i < n; ++i) out[i] = rhs[i];
    for (int k = 0; k < n; ++k)
        if (f->piv[k] != k) {
            double t = out[k];
            out[k] = out[f->piv[k]];
            out[f->piv[k]] = t;
        }
    for (int i = 0; i < n; ++i) {
        double s = out[i];
        for (int j = 0; j < i; ++j) s -= a[(size_t)i * n + j] * out[j];
        out[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = out[i];
        for (int j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * out[j];
        out[i] = s / a[(size_t)i * n + i];
    }
}

static void lu_free(orc_lu *f)
{
    free(f->lu);
    free(f->piv);
    f->lu = NULL;
    f->piv = NULL;
}

/* ctypes entry: solve the CSC system A x = g with the dense LU. */
int orc_dense_solve_csc(int n, const int *Ap, const int *Ai, const double *Ax, const double *g, double *xout)
{
    orc_lu f;
    int rc = lu_factor_csc(&f, n, Ap, Ai, Ax);
    if (rc == 0) lu_solve(&f, g, xout);
    lu_free(&f);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* The simulation driver: integrate / calculateG / stages / run             */
/* (simulation.cpp:62-84, 133-193, 417-650).                                 */
/* ---------------------------------------------------------------------- */
enum { ORC_ORIGINAL = 0, ORC_END_BIG = 1, ORC_END_FIRST_SMALL = 2, ORC_END_SECOND_SMALL = 3 };
enum { ORC_ZERO = 0, ORC_FIXED_RATE = 1, ORC_SPRING = 2, ORC_CONSTANT = 3 };

/* The run-loop settings of SimulationData (simulation_data.h:84-254) that
 * Simulation::run and stage III read (simulation.cpp:417-433, 553-633). */
typedef struct {
    int isTimeLimit;
    double timeLimit;
    int isStrainIncreaseLimit;
    double strainIncreaseLimit;
    int isStepCountLimit;
    long stepCountLimit;
    int countAvalanches;
    long avalancheTriggerLimit;
    double avalancheSpeedThreshold;
    int isSpeedLimit;
    double speedLimit;
    int isStressLimit;
    double stressLimit;
    int isSpeedThresholdForCutoffChange;
    double speedThresholdForCutoffChange;
    int isSaveSubConfigs;
    long subConfigDelay, subConfigDelayDuringAvalanche;
    long remainingFinalSteps;
} orc_run_opts;

typedef struct {
    int n, pc;
    double *x, *y, *b;                       /* dislocations (state)        */
    double *bigx, *fsx, *ssx;                /* stage positions (x only)    */
    double *px, *py;
    double A, KASQR;
    double cutOffMultiplier;
    orc_cut cut;
    double prec;
    int ic;
    double stepSize, simTime;
    /* stress protocol (stress_protocol.cpp, fixed_rate_protocol.cpp, spring_protocol.cpp) */
    int protocol;
    double rate, springConstant, constStress;
    double stressValues[4];
    int currentStressStateType;
    double totalAccumulatedStrainIncrease;
    int calculateStrain;
    int orderParameterOn;
    /* speeds */
    double *initSpeed, *initSpeed2, *speed, *speed2, *g, *dVec, *delta;
    /* CSC */
    long cap;
    int *Ap, *Ai, *indexes;
    double *Ax;
    orc_lu lu;
    orc_ph ph;
    /* Simulation members (simulation.h:69-81) */
    int succesfulStep, initSpeedCalculationIsNeeded, firstStepRequest;
    double energy, energyAccum, vsquare, vsquare1, vsquare2;
    long succesfulSteps, failedSteps;
    int isMaxStepSizeLimit;
    double maxStepSizeLimit;
    double sumAvgSpeed;
    /* log: the 12 columns of the reference's log lines, one row per line */
    double *logrows;
    long logcap, loglen;
    double *ysave;
    /* run loop (simulation.cpp:417-433, 553-633) */
    orc_run_opts run;
    long avalancheCount;
    int inAvalanche, finish, inFinal;
    long subconfigDistanceCounter;
    /* sub-configurations written: simulation time and positions of each */
    double *subTimes, *subX;
    long subLen, subCap;
} orc_sim;

static double protocol_get(const orc_sim *s, int type)
{
    if (s->protocol == ORC_ZERO) return 0;
    if (s->protocol == ORC_CONSTANT) return s->constStress;
    return s->stressValues[type];
}

static void protocol_calc(orc_sim *s, double t, int type, double strain)
{
    if (s->protocol == ORC_FIXED_RATE) {
        s->stressValues[type] = t * s->rate;
    } else if (s->protocol == ORC_SPRING) {
        double value = s->springConstant * (t * s->rate - strain);
        if (value < 0) value = 0;
        s->stressValues[type] = value;
    }
}

/* utility.h:92-100 */
static double strain_increment(int n, const double *b, const double *oldx, const double *newx)
{
    double r = 0;
    for (int i = 0; i < n; ++i) r += b[i] * (newx[i] - oldx[i]);
    return r;
}

static void sim_speeds(orc_sim *s, const double *xs, double *res)
{
    orc_sys sys = {s->n, s->pc, xs, s->y, s->b, s->px, s->py, s->A, s->KASQR};
    speeds_impl(&sys, protocol_get(s, s->currentStressStateType), res, &s->ph);
}

/* simulation.cpp:133-193 */
static void calculate_g(orc_sim *s, double stepsize, const double *newx, const double *oldx, int useSpeed2,
                        int calculateInitSpeed, int useInitSpeedForFirstStep, int origin, int end)
{
    double *isp = s->initSpeed, *csp = s->speed;
    if (useSpeed2) {
        isp = s->initSpeed2;
        csp = s->speed2;
    }
    if (calculateInitSpeed) {
        double t = s->simTime;
        double tasi = s->totalAccumulatedStrainIncrease;
        if (origin == ORC_END_FIRST_SMALL) {
            t += s->stepSize * 0.5;
            tasi += strain_increment(s->n, s->b, s->x, s->fsx);
        }
        protocol_calc(s, t, origin, tasi);
        s->currentStressStateType = origin;
        sim_speeds(s, oldx, isp);
    }
    if (useInitSpeedForFirstStep) {
        csp = isp;
    } else {
        double t = s->simTime + s->stepSize;
        double tasi = s->totalAccumulatedStrainIncrease;
        if (end == ORC_END_FIRST_SMALL) {
            t -= s->stepSize * 0.5;
            tasi += strain_increment(s->n, s->b, s->x, s->fsx);
        } else if (end == ORC_END_SECOND_SMALL) {
            tasi += strain_increment(s->n, s->b, s->x, s->fsx) + strain_increment(s->n, s->b, s->fsx, s->ssx);
        } else if (end == ORC_END_BIG) {
            tasi += strain_increment(s->n, s->b, s->x, s->bigx);
        }
        protocol_calc(s, t, end, tasi);
        s->currentStressStateType = end;
        sim_speeds(s, newx, csp);
    }
    for (int i = 0; i < s->n; ++i)
        s->g[i] = newx[i] - (1.0 + s->dVec[i]) * 0.5 * stepsize * csp[i] - oldx[i] -
                  (1.0 - s->dVec[i]) * 0.5 * stepsize * isp[i];
}

/* simulation.cpp:62-84 */
static int integrate(orc_sim *s, double stepsize, double *newx, const double *oldx, int useSpeed2,
                     int calculateInitSpeed, int origin, int end)
{
    orc_sys sys = {s->n, s->pc, newx, s->y, s->b, s->px, s->py, s->A, s->KASQR};
    long nnz = jac_pass1(&sys, &s->cut, stepsize, s->cap, s->Ap, s->Ai, s->Ax);
    if (nnz < 0) return -1;
    jac_pass23(s->n, s->Ap, s->Ai, s->Ax, s->dVec, s->indexes);
    if (lu_factor_csc(&s->lu, s->n, s->Ap, s->Ai, s->Ax) != 0) return -2;
    for (int i = 0; i < s->ic; ++i) {
        if (i > 0)
            calculate_g(s, stepsize, newx, oldx, useSpeed2, 0, 0, origin, end);
        else
            calculate_g(s, stepsize, newx, oldx, useSpeed2, calculateInitSpeed, s->protocol == ORC_ZERO, origin, end);
        lu_solve(&s->lu, s->g, s->delta);
        for (int j = 0; j < s->n; ++j) newx[j] -= s->delta[j];
    }
    lu_free(&s->lu);
    return 0;
}

static double sum_abs(const double *v, int n)
{
    double a = 0;
    for (int i = 0; i < n; ++i) a = a + fabs(v[i]);
    return a;
}

static double sum_sq(const double *v, int n)
{
    double a = 0;
    for (int i = 0; i < n; ++i) a = a + v[i] * v[i];
    return a;
}

static void log_row(orc_sim *s, const double row[12])
{
    if (s->loglen >= s->logcap) {
        long nc = s->logcap ? s->logcap * 2 : 64;
        double *r = (double *)realloc(s->logrows, (size_t)nc * 12 * sizeof(double));
        if (!r) return;
        s->logrows = r;
        s->logcap = nc;
    }
    memcpy(s->logrows + s->loglen * 12, row, 12 * sizeof(double));
    s->loglen++;
}

static void sub_record(orc_sim *s)
{
    if (s->subLen >= s->subCap) {
        long nc = s->subCap ? s->subCap * 2 : 16;
        double *tt = (double *)realloc(s->subTimes, (size_t)nc * sizeof(double));
        if (!tt) return;
        s->subTimes = tt;
        double *xx = (double *)realloc(s->subX, (size_t)nc * (size_t)s->n * sizeof(double));
        if (!xx) return;
        s->subX = xx;
        s->subCap = nc;
    }
    s->subTimes[s->subLen] = s->simTime;
    memcpy(s->subX + (size_t)s->subLen * (size_t)s->n, s->x, (size_t)s->n * sizeof(double));
    s->subLen++;
}

/* simulation.cpp:435-650 ; returns 1 on a successful step, 0 on a failed
 * step, <0 on error.  Log column 9 (wall time) is recorded as 0. */
int orc_sim_step(orc_sim *s)
{
    const int n = s->n;
    /* Stage I (443-492) */
    s->energyAccum = 0;
    double sumAvgSp = 0;
    s->vsquare = 0;
    s->currentStressStateType = ORC_ORIGINAL;
    if (s->firstStepRequest) {
        protocol_calc(s, s->simTime, ORC_ORIGINAL, s->totalAccumulatedStrainIncrease);
        sim_speeds(s, s->x, s->initSpeed);
        s->initSpeedCalculationIsNeeded = 0;
        sumAvgSp = sum_abs(s->initSpeed, n) / (double)n;
        s->vsquare = sum_sq(s->initSpeed, n);
        double row[12] = {s->simTime, (double)s->succesfulSteps, (double)s->failedSteps, 0, sumAvgSp,
                          s->cut.cutOff, NAN, protocol_get(s, s->currentStressStateType), NAN,
                          s->totalAccumulatedStrainIncrease, s->vsquare, s->energy};
        log_row(s, row);
        s->firstStepRequest = 0;
    }
    memcpy(s->bigx, s->x, (size_t)n * sizeof(double));
    if (integrate(s, s->stepSize, s->bigx, s->x, 0, s->initSpeedCalculationIsNeeded, ORC_ORIGINAL, ORC_END_BIG)) return -1;
    s->succesfulStep = 0;
    /* Stage II (494-499) */
    memcpy(s->fsx, s->x, (size_t)n * sizeof(double));
    if (integrate(s, 0.5 * s->stepSize, s->fsx, s->x, 0, 0, ORC_ORIGINAL, ORC_END_FIRST_SMALL)) return -1;
    /* Stage III (501-650) */
    s->sumAvgSpeed = 0;
    memcpy(s->ssx, s->fsx, (size_t)n * sizeof(double));
    if (integrate(s, 0.5 * s->stepSize, s->ssx, s->fsx, 1, 1, ORC_END_FIRST_SMALL, ORC_END_SECOND_SMALL)) return -1;
    s->vsquare1 = sum_sq(s->initSpeed, n);
    s->vsquare2 = sum_sq(s->initSpeed2, n);
    s->energyAccum = (s->vsquare1 + s->vsquare2) * 0.5 * s->stepSize * 0.5;
    for (int i = 0; i < n; ++i) ph_update_error(&s->ph, fabs(s->bigx[i] - s->ssx[i]), (unsigned)i);
    if (s->ph.maxErrorRatioSqr < 1.0) {
        s->succesfulStep = 1;
        s->initSpeedCalculationIsNeeded = 1;
        if (s->inFinal) s->run.remainingFinalSteps--;
        if (s->calculateStrain) {
            s->totalAccumulatedStrainIncrease += strain_increment(n, s->b, s->x, s->fsx);
            s->totalAccumulatedStrainIncrease += strain_increment(n, s->b, s->fsx, s->ssx);
        }
        double *t = s->x;
        s->x = s->ssx;
        s->ssx = t;
        s->simTime += s->stepSize;
        s->succesfulSteps++;
        /* simulation.cpp:407-415, 535-539: sum of b_i * speed_i (speed, not initSpeed) */
        double orderParameter = NAN;
        if (s->orderParameterOn) {
            orderParameter = 0;
            for (int i = 0; i < n; ++i) orderParameter += s->b[i] * s->speed[i];
        }
        s->currentStressStateType = ORC_ORIGINAL;
        protocol_calc(s, s->simTime, ORC_ORIGINAL, s->totalAccumulatedStrainIncrease);
        sim_speeds(s, s->x, s->initSpeed);
        s->initSpeedCalculationIsNeeded = 0;
        s->sumAvgSpeed = sum_abs(s->initSpeed, n) / (double)n;
        /* simulation.cpp:553-555 */
        if (s->sumAvgSpeed < s->run.speedLimit && s->run.isSpeedLimit) s->finish = 1;
        s->vsquare = sum_sq(s->initSpeed, n);
        /* simulation.cpp:559-570 */
        if (s->run.countAvalanches) {
            if (s->inAvalanche && s->sumAvgSpeed < s->run.avalancheSpeedThreshold) {
                s->avalancheCount++;
                s->inAvalanche = 0;
            } else if (s->sumAvgSpeed > s->run.avalancheSpeedThreshold) {
                s->inAvalanche = 1;
            }
        }
        s->energyAccum += (s->vsquare2 + s->vsquare) * 0.5 * s->stepSize * 0.5;
        const double loggedCutOff = s->cut.cutOff;
        /* simulation.cpp:603-611 */
        if (s->run.isStressLimit && protocol_get(s, ORC_ORIGINAL) > s->run.stressLimit) s->finish = 1;
        if (s->run.isSpeedThresholdForCutoffChange && s->run.speedThresholdForCutoffChange > s->sumAvgSpeed) {
            s->cutOffMultiplier = 1e20;
            orc_cutoff(n, s->cutOffMultiplier, &s->cut.cutOff, &s->cut.cutOffSqr, &s->cut.onePerCutOffSqr);
        }
        s->energy += s->energyAccum;
        double row[12] = {s->simTime, (double)s->succesfulSteps, (double)s->failedSteps, s->ph.maxErrorRatioSqr,
                          s->sumAvgSpeed, loggedCutOff, orderParameter, protocol_get(s, ORC_ORIGINAL), 0.0,
                          s->calculateStrain ? s->totalAccumulatedStrainIncrease : NAN, s->vsquare, s->energy};
        log_row(s, row);
        /* simulation.cpp:619-633: the file is named by the simulation time */
        if (s->run.isSaveSubConfigs) {
            if ((!s->inAvalanche && s->run.subConfigDelay >= s->subconfigDistanceCounter) ||
                (s->inAvalanche && s->run.subConfigDelayDuringAvalanche >= s->subconfigDistanceCounter)) {
                s->subconfigDistanceCounter = 0;
                sub_record(s);
            } else {
                s->subconfigDistanceCounter++;
            }
        }
    } else {
        s->failedSteps++;
    }
    s->stepSize = orc_new_step_size(s->ph.maxErrorRatioSqr, s->stepSize);
    ph_reset(&s->ph);
    if (s->isMaxStepSizeLimit && s->maxStepSizeLimit < s->stepSize) s->stepSize = s->maxStepSizeLimit;
    return s->succesfulStep;
}

/* Create a simulation.  protocol: 0 zero, 1 fixed-rate (p1=rate), 2 spring
 * (p1=rate, p2=spring constant), 3 constant (p1=stress). */
orc_sim *orc_sim_create(int n, const double *x, const double *y, const double *b, int pc, const double *px,
                        const double *py, double A, double KASQR, double cutOffMultiplier, double prec,
                        double stepSize, int protocol, double p1, double p2, int calculateStrain, long cap)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->n = n;
    s->pc = pc;
    size_t nb = (size_t)n * sizeof(double);
    s->x = (double *)malloc(nb);
    s->y = (double *)malloc(nb);
    s->b = (double *)malloc(nb);
    s->bigx = (double *)malloc(nb);
    s->fsx = (double *)malloc(nb);
    s->ssx = (double *)malloc(nb);
    memcpy(s->x, x, nb);
    memcpy(s->y, y, nb);
    memcpy(s->b, b, nb);
    s->px = (double *)malloc((size_t)(pc > 0 ? pc : 1) * sizeof(double));
    s->py = (double *)malloc((size_t)(pc > 0 ? pc : 1) * sizeof(double));
    if (pc > 0) {
        memcpy(s->px, px, (size_t)pc * sizeof(double));
        memcpy(s->py, py, (size_t)pc * sizeof(double));
    }
    s->A = A;
    s->KASQR = KASQR;
    s->cutOffMultiplier = cutOffMultiplier;
    orc_cutoff(n, cutOffMultiplier, &s->cut.cutOff, &s->cut.cutOffSqr, &s->cut.onePerCutOffSqr);
    s->prec = prec;
    s->ic = 2;
    s->stepSize = stepSize;
    s->protocol = protocol;
    if (protocol == ORC_FIXED_RATE || protocol == ORC_SPRING) s->rate = p1;
    if (protocol == ORC_SPRING) s->springConstant = p2;
    if (protocol == ORC_CONSTANT) s->constStress = p1;
    s->calculateStrain = calculateStrain;
    s->initSpeed = (double *)calloc((size_t)n, sizeof(double));
    s->initSpeed2 = (double *)calloc((size_t)n, sizeof(double));
    s->speed = (double *)calloc((size_t)n, sizeof(double));
    s->speed2 = (double *)calloc((size_t)n, sizeof(double));
    s->g = (double *)calloc((size_t)n, sizeof(double));
    s->dVec = (double *)calloc((size_t)n, sizeof(double));
    s->delta = (double *)calloc((size_t)n, sizeof(double));
    s->cap = cap > 0 ? cap : (long)n * n;
    s->Ap = (int *)calloc((size_t)n + 1, sizeof(int));
    s->Ai = (int *)calloc((size_t)s->cap, sizeof(int));
    s->Ax = (double *)calloc((size_t)s->cap, sizeof(double));
    s->indexes = (int *)calloc((size_t)n, sizeof(int));
    /* Simulation ctor (simulation.cpp:38-55): setMinPrecisity then setSize */
    s->ph.n = n;
    s->ph.minPrec = prec;
    s->ph.minPrecSqr = prec * prec;
    s->ph.tol = (double *)malloc(nb);
    s->ph.err = (double *)malloc(nb);
    for (int i = 0; i < n; ++i) {
        s->ph.tol[i] = s->ph.minPrecSqr;
        s->ph.err[i] = 0;
    }
    s->succesfulStep = 1;
    s->initSpeedCalculationIsNeeded = 1;
    s->firstStepRequest = 1;
    return s;
}

void orc_sim_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->x); free(s->y); free(s->b); free(s->bigx); free(s->fsx); free(s->ssx);
    free(s->px); free(s->py);
    free(s->initSpeed); free(s->initSpeed2); free(s->speed); free(s->speed2);
    free(s->g); free(s->dVec); free(s->delta);
    free(s->Ap); free(s->Ai); free(s->Ax); free(s->indexes);
    free(s->ph.tol); free(s->ph.err);
    free(s->logrows);
    free(s->subTimes); free(s->subX);
    free(s);
}

void orc_sim_set_step_size(orc_sim *s, double h) { s->stepSize = h; }
void orc_sim_set_order_parameter(orc_sim *s, int on) { s->orderParameterOn = on; }
void orc_sim_set_max_step_size(orc_sim *s, double hmax)
{
    s->isMaxStepSizeLimit = 1;
    s->maxStepSizeLimit = hmax;
}
double orc_sim_get_step_size(const orc_sim *s) { return s->stepSize; }
double orc_sim_get_time(const orc_sim *s) { return s->simTime; }
long orc_sim_succesful_steps(const orc_sim *s) { return s->succesfulSteps; }
long orc_sim_failed_steps(const orc_sim *s) { return s->failedSteps; }
double orc_sim_energy(const orc_sim *s) { return s->energy; }
double orc_sim_strain(const orc_sim *s) { return s->totalAccumulatedStrainIncrease; }
double orc_sim_max_error_ratio_last(const orc_sim *s) { return s->ph.maxErrorRatioSqr; }
void orc_sim_get_x(const orc_sim *s, double *out) { memcpy(out, s->x, (size_t)s->n * sizeof(double)); }
long orc_sim_log_rows(const orc_sim *s) { return s->loglen; }
void orc_sim_get_log(const orc_sim *s, double *out) { memcpy(out, s->logrows, (size_t)s->loglen * 12 * sizeof(double)); }
void orc_sim_get_speeds(const orc_sim *s, double *initSpeed, double *initSpeed2)
{
    if (initSpeed) memcpy(initSpeed, s->initSpeed, (size_t)s->n * sizeof(double));
    if (initSpeed2) memcpy(initSpeed2, s->initSpeed2, (size_t)s->n * sizeof(double));
}

/* Stage I positions of the last step (the big step, simulation.cpp:482-488)
 * and stage II's (the first small step, 494-499). */
void orc_sim_get_big_x(const orc_sim *s, double *out) { memcpy(out, s->bigx, (size_t)s->n * sizeof(double)); }
void orc_sim_get_fs_x(const orc_sim *s, double *out) { memcpy(out, s->fsx, (size_t)s->n * sizeof(double)); }

/* Entry points for the mathematical pins of the stepping path (tests only).
 * orc_sim_residual: one calculateG (simulation.cpp:133-193) of the stage-I
 * kind at the given new/old positions, with dVec and the old speeds given
 * (so the residual can be differentiated with them held fixed); the new
 * speeds are computed at newx as the stepping path computes them. */
void orc_sim_residual(orc_sim *s, double stepsize, const double *newx, const double *oldx, const double *dVec,
                      const double *v_old, double *g_out)
{
    memcpy(s->dVec, dVec, (size_t)s->n * sizeof(double));
    memcpy(s->initSpeed, v_old, (size_t)s->n * sizeof(double));
    calculate_g(s, stepsize, newx, oldx, 0, 0, 0, ORC_ORIGINAL, ORC_END_BIG);
    memcpy(g_out, s->g, (size_t)s->n * sizeof(double));
}

/* orc_sim_stage: one integrate (simulation.cpp:62-84) of stage-I kind from
 * the simulation's current positions with step size h: newx receives the
 * end positions, dVec_out (nullable) the d vector of the assembly. */
/* Simulation::run (simulation.cpp:417-433) with the given settings: step
 * while every active limit allows, then the post-relax steps.  Returns the
 * number of attempts made, <0 on error. */
long orc_sim_run(orc_sim *s, const orc_run_opts *o)
{
    s->run = *o;
    long attempts = 0;
    while (((s->run.isTimeLimit && s->simTime < s->run.timeLimit) || !s->run.isTimeLimit) &&
           ((s->run.isStrainIncreaseLimit && s->totalAccumulatedStrainIncrease < s->run.strainIncreaseLimit) ||
            !s->run.isStrainIncreaseLimit) &&
           ((s->run.isStepCountLimit && s->succesfulSteps < s->run.stepCountLimit) || !s->run.isStepCountLimit) &&
           ((s->run.countAvalanches && s->avalancheCount < s->run.avalancheTriggerLimit) || !s->run.countAvalanches) &&
           !s->finish) {
        if (orc_sim_step(s) < 0) return -1;
        attempts++;
    }
    while (s->run.remainingFinalSteps > 0) {
        s->inFinal = 1;
        if (orc_sim_step(s) < 0) return -1;
        attempts++;
    }
    return attempts;
}

/* Run-loop state: [avalancheCount, inAvalanche, finish, remainingFinalSteps,
 * cutOff, cutOffMultiplier, subconfigDistanceCounter]. */
void orc_sim_run_state(const orc_sim *s, double *out)
{
    out[0] = (double)s->avalancheCount;
    out[1] = s->inAvalanche;
    out[2] = s->finish;
    out[3] = (double)s->run.remainingFinalSteps;
    out[4] = s->cut.cutOff;
    out[5] = s->cutOffMultiplier;
    out[6] = (double)s->subconfigDistanceCounter;
}

/* Sub-configurations recorded so far: times (and positions, n per record)
 * into the caller's arrays of cap records; returns the count. */
long orc_sim_subconfigs(const orc_sim *s, double *times, double *xs, long cap)
{
    for (long k = 0; k < s->subLen && k < cap; ++k) {
        if (times) times[k] = s->subTimes[k];
        if (xs) memcpy(xs + (size_t)k * (size_t)s->n, s->subX + (size_t)k * (size_t)s->n, (size_t)s->n * sizeof(double));
    }
    return s->subLen;
}

int orc_sim_stage(orc_sim *s, double h, double *newx, double *dVec_out)
{
    memcpy(newx, s->x, (size_t)s->n * sizeof(double));
    int rc = integrate(s, h, newx, s->x, 0, 1, ORC_ORIGINAL, ORC_END_BIG);
    if (dVec_out) memcpy(dVec_out, s->dVec, (size_t)s->n * sizeof(double));
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Batched evaluations for the tests                                        */
/* ---------------------------------------------------------------------- */

/* Host C library over an array: fn 0 cos, 1 sin, 2 exp, 3 expm1, 4 cosh,
 * 5 sinh, 6 sqrt, 7 pow(x, 1./3.), 8 normalize, 9 pow(M_E, x),
 * 10 pow(x, -2). */
void orc_libm_batch(int fn, long n, const double *x, double *out)
{
    for (long i = 0; i < n; ++i) {
        double v = x[i], r = 0;
        switch (fn) {
        case 0: r = cos(v); break;
        case 1: r = sin(v); break;
        case 2: r = exp(v); break;
        case 3: r = expm1(v); break;
        case 4: r = cosh(v); break;
        case 5: r = sinh(v); break;
        case 6: r = sqrt(v); break;
        case 7: r = pow(v, 1. / 3.); break;
        case 8: r = orc_normalize(v); break;
        case 9: r = pow(M_E, v); break;
        case 10: r = pow(v, -0.2e1); break;
        default: r = 0; break;
        }
        out[i] = r;
    }
}

void orc_xy_batch(long n, const double *dx, const double *dy, double *out, double *out_dx)
{
    for (long i = 0; i < n; ++i) {
        if (out) out[i] = orc_xy(dx[i], dy[i]);
        if (out_dx) out_dx[i] = orc_xy_diff_x(dx[i], dy[i]);
    }
}

/* Per-target sum of |pair term| (the scale of the rounding of any ordering of
 * the pair sum), for conditioning-aware speed comparisons. */
void orc_speeds_abs(int n, const double *x, const double *y, const double *b, double *S)
{
    for (int i = 0; i < n; ++i) S[i] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double t = fabs(b[i] * b[j] * orc_xy(orc_normalize(x[i] - x[j]), orc_normalize(y[i] - y[j])));
            S[i] += t;
            S[j] += t;
        }
}
