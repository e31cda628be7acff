// External stress protocols (reference: include/StressProtocols/stress_protocol.h:31-82).
#ifndef SDDDST_CORE_STRESS_PROTOCOL_H
#define SDDDST_CORE_STRESS_PROTOCOL_H

#include <string>
#include <vector>

#include "../dislocation.h"

namespace sdddstCore {

enum StressProtocolStepType { Original, EndOfBigStep, EndOfFirstSmallStep, EndOfSecondSmallStep };

// Zero external stress ("zero-stress").
class StressProtocol {
public:
    StressProtocol();
    virtual ~StressProtocol();

    virtual void calculateStress(double simulationTime, const std::vector<Dislocation> &dislocations,
                                 StressProtocolStepType type);
    virtual double getStress(StressProtocolStepType type);
    virtual double getStressDerivative(double simulationTime);
    virtual bool getTriggerState();
    virtual std::string getType();
};

}  // namespace sdddstCore

#endif
