// Analytic periodic shear-stress field (reference: include/Fields/AnalyticField.h:30-39).
// Simulation evaluates it on the device; xy() / xy_diff_x() on the host
// evaluate the same device implementation for single points.
#ifndef SDDDST_CORE_ANALYTIC_FIELD_H
#define SDDDST_CORE_ANALYTIC_FIELD_H

#include "../constants.h"
#include "Field.h"

namespace sdddstCore {

class AnalyticField : public Field {
public:
    AnalyticField();
    virtual ~AnalyticField();

    virtual double xy(double dx, double dy);
    virtual double xy_diff_x(double dx, double dy);
    // Second derivative: used only by the eigenvalue analysis, which this
    // library does not provide; returns NaN.
    virtual double xy_diff_x2(double dx, double dy);

    virtual int deviceKind() const { return 1; }
};

}  // namespace sdddstCore

#endif
