// Stress protocols (reference: src/StressProtocols/stress_protocol.cpp:32-57,
// fixed_rate_protocol.cpp:39-83, spring_protocol.cpp:33-59) and the
// constant protocol of this build.
#include "../../include/sdddst/StressProtocols/constant_stress_protocol.h"
#include "../../include/sdddst/StressProtocols/fixed_rate_protocol.h"
#include "../../include/sdddst/StressProtocols/spring_protocol.h"
#include "../../include/sdddst/StressProtocols/stress_protocol.h"

using namespace sdddstCore;

StressProtocol::StressProtocol() {}
StressProtocol::~StressProtocol() {}
void StressProtocol::calculateStress(double, const std::vector<Dislocation> &, StressProtocolStepType) {}
double StressProtocol::getStress(StressProtocolStepType) { return 0; }
double StressProtocol::getStressDerivative(double) { return 0; }
bool StressProtocol::getTriggerState() { return false; }
std::string StressProtocol::getType() { return "zero-stress"; }

FixedRateProtocol::FixedRateProtocol() : StressProtocol(), rate(0), stressValues{0, 0, 0, 0} {}
FixedRateProtocol::~FixedRateProtocol() {}

void FixedRateProtocol::calculateStress(double simulationTime, const std::vector<Dislocation> &,
                                        StressProtocolStepType type)
{
    stressValues[type] = simulationTime * rate;
}

double FixedRateProtocol::getStress(StressProtocolStepType type) { return stressValues[type]; }
double FixedRateProtocol::getStressDerivative(double) { return rate; }
std::string FixedRateProtocol::getType() { return "fixed-rate-stress"; }
double FixedRateProtocol::getRate() const { return rate; }
void FixedRateProtocol::setRate(double value) { rate = value; }

SpringProtocol::SpringProtocol() : FixedRateProtocol(), springConstant(1.0) {}
SpringProtocol::~SpringProtocol() {}

void SpringProtocol::calculateStress(double simulationTime, const std::vector<Dislocation> &,
                                     StressProtocolStepType type, double strain)
{
    double value = springConstant * (simulationTime * rate - strain);
    if (value < 0) value = 0;
    stressValues[type] = value;
}

std::string SpringProtocol::getType() { return "spring-stress"; }
double SpringProtocol::getSpringConstant() const { return springConstant; }
void SpringProtocol::setSpringConstant(double v) { springConstant = v; }

ConstantStressProtocol::ConstantStressProtocol(double stress) : StressProtocol(), value(stress) {}
ConstantStressProtocol::~ConstantStressProtocol() {}
double ConstantStressProtocol::getStress(StressProtocolStepType) { return value; }
std::string ConstantStressProtocol::getType() { return "constant-stress"; }
double ConstantStressProtocol::getValue() const { return value; }
void ConstantStressProtocol::setValue(double v) { value = v; }
