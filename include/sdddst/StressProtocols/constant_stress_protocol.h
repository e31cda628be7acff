// Constant external stress sigma ("constant-stress").  The reference has no
// such protocol; it is the time-independent case of FixedRateProtocol's
// interface.  Because the stress does not depend on time, the first Newton
// iteration may reuse the initial speeds exactly as "zero-stress" does
// (simulation.cpp:75, 163-166): the new positions equal the old ones there,
// so the skipped evaluation would return the same bits.
#ifndef SDDDST_CORE_CONSTANT_STRESS_PROTOCOL_H
#define SDDDST_CORE_CONSTANT_STRESS_PROTOCOL_H

#include "stress_protocol.h"

namespace sdddstCore {

class ConstantStressProtocol : public StressProtocol {
public:
    explicit ConstantStressProtocol(double stress = 0.0);
    virtual ~ConstantStressProtocol();

    virtual double getStress(StressProtocolStepType type);
    virtual std::string getType();

    double getValue() const;
    void setValue(double v);

protected:
    double value;
};

}  // namespace sdddstCore

#endif
