// Dislocation stress field interface (reference: include/Fields/Field.h:28-37).
#ifndef SDDDST_CORE_FIELD_H
#define SDDDST_CORE_FIELD_H

namespace sdddstCore {

class Field {
public:
    Field();
    virtual ~Field();

    virtual double xy(double dx, double dy);
    virtual double xy_diff_x(double dx, double dy);
    virtual double xy_diff_x2(double dx, double dy);

    // Field kind the device kernels implement (0: none; 1: analytic periodic).
    virtual int deviceKind() const { return 0; }
};

}  // namespace sdddstCore

#endif
