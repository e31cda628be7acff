// sigma = k * max(0, rate * t - strain) (reference: include/StressProtocols/spring_protocol.h).
#ifndef SDDDST_SPRING_PROTOCOL_H
#define SDDDST_SPRING_PROTOCOL_H

#include "fixed_rate_protocol.h"

namespace sdddstCore {

class SpringProtocol : public FixedRateProtocol {
public:
    SpringProtocol();
    virtual ~SpringProtocol();

    using FixedRateProtocol::calculateStress;
    virtual void calculateStress(double simulationTime, const std::vector<Dislocation> &dislocations,
                                 StressProtocolStepType type, double strain);
    virtual std::string getType();

    double getSpringConstant() const;
    void setSpringConstant(double newSpringConstant);

protected:
    double springConstant;
};

}  // namespace sdddstCore

#endif
