// Test-only harness around the reference's own translation units
// (AnalyticField.cpp, Field.cpp, precision_handler.cpp and the stress
// protocols), built by oracle/Makefile into oracle/_ref/libsdddst_ref.so from
// the sources where they lie.  It exposes plain C entry points so the tests
// can compare the oracle with the reference's own per-pair field functions
// and precision handler, bit for bit.  simulation.cpp itself needs umfpack.h,
// which this image lacks, so its loops are not built here; ref_speeds below
// drives the reference's AnalyticField::xy through the Field virtual with the
// loop order of simulation.cpp:86-131.
#include "Fields/AnalyticField.h"
#include "Fields/Field.h"
#include "precision_handler.h"
#include "StressProtocols/fixed_rate_protocol.h"
#include "StressProtocols/spring_protocol.h"
#include "StressProtocols/stress_protocol.h"
#include "utility.h"

#include <memory>
#include <vector>

using namespace sdddstCore;

extern "C" {

double ref_normalize(double v)
{
    normalize(v);
    return v;
}

double ref_xy(double dx, double dy)
{
    AnalyticField f;
    Field *p = &f;
    return p->xy(dx, dy);
}

double ref_xy_diff_x(double dx, double dy)
{
    AnalyticField f;
    Field *p = &f;
    return p->xy_diff_x(dx, dy);
}

double ref_xy_diff_x2(double dx, double dy)
{
    AnalyticField f;
    Field *p = &f;
    return p->xy_diff_x2(dx, dy);
}

void ref_xy_batch(long n, const double *dx, const double *dy, double *out, double *out_dx)
{
    AnalyticField f;
    Field *p = &f;
    for (long i = 0; i < n; ++i) {
        if (out) out[i] = p->xy(dx[i], dy[i]);
        if (out_dx) out_dx[i] = p->xy_diff_x(dx[i], dy[i]);
    }
}

// simulation.cpp:86-131 without point defects; tolerances through the
// reference's PrecisionHandler (tol_out receives toleranceAndError.first via
// updateError probes is not possible, so the handler is exercised separately).
void ref_speeds(int n, const double *x, const double *y, const double *b, double sigma, double *res)
{
    AnalyticField f;
    Field *tau = &f;
    for (int i = 0; i < n; ++i) res[i] = 0;
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            double dx = x[i] - x[j];
            normalize(dx);
            double dy = y[i] - y[j];
            normalize(dy);
            double tmp = b[i] * b[j] * tau->xy(dx, dy);
            res[i] += tmp;
            res[j] -= tmp;
        }
        res[i] += b[i] * sigma;
    }
}

// PrecisionHandler round trip: prec, then nupd tolerance updates (r2, id),
// then nerr error updates (err, id); returns maxErrorRatioSqr and the new
// step size for oldStep.
double ref_precision_sequence(unsigned size, double prec, int nupd, const double *r2, const unsigned *ids,
                              int nerr, const double *err, const unsigned *eids, double oldStep, double *newStep)
{
    PrecisionHandler ph;
    ph.setMinPrecisity(prec);
    ph.setSize(size);
    for (int k = 0; k < nupd; ++k) ph.updateTolerance(r2[k], ids[k]);
    for (int k = 0; k < nerr; ++k) ph.updateError(err[k], eids[k]);
    if (newStep) *newStep = ph.getNewStepSize(oldStep);
    return ph.getMaxErrorRatioSqr();
}

// Stress protocols: value reported by getStress(type) after calculateStress.
double ref_protocol_stress(int kind, double rate, double spring, double t, double strain, int type)
{
    std::vector<Dislocation> none;
    StressProtocolStepType st = static_cast<StressProtocolStepType>(type);
    if (kind == 0) {
        StressProtocol p;
        p.calculateStress(t, none, st);
        return p.getStress(st);
    }
    if (kind == 1) {
        FixedRateProtocol p;
        p.setRate(rate);
        p.calculateStress(t, none, st);
        return p.getStress(st);
    }
    SpringProtocol p;
    p.setRate(rate);
    p.setSpringConstant(spring);
    p.calculateStress(t, none, st, strain);
    return p.getStress(st);
}

}  // extern "C"
