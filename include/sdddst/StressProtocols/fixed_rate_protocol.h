// sigma = rate * t (reference: include/StressProtocols/fixed_rate_protocol.h).
#ifndef SDDDST_CORE_FIXED_RATE_PROTOCOL_H
#define SDDDST_CORE_FIXED_RATE_PROTOCOL_H

#include "stress_protocol.h"

namespace sdddstCore {

class FixedRateProtocol : public StressProtocol {
public:
    FixedRateProtocol();
    virtual ~FixedRateProtocol();

    virtual void calculateStress(double simulationTime, const std::vector<Dislocation> &dislocations,
                                 StressProtocolStepType type);
    virtual double getStress(StressProtocolStepType type);
    virtual double getStressDerivative(double simulationTime);
    virtual std::string getType();

    double getRate() const;
    void setRate(double value);

protected:
    double rate;
    double stressValues[4];
};

}  // namespace sdddstCore

#endif
