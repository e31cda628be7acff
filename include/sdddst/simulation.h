// Implicit-step integrator (reference: include/simulation.h:38-86,
// src/simulation.cpp).  Same public methods; the pair speeds, the Jacobian,
// the residual, the Newton update and the error controller run on the
// MI355X through the C ABI of include/sdddst_amd.h.  The linear solve is
// handed to host UMFPACK exactly as in the reference when libumfpack is
// available, or done on the device (see LinearSolver).
#ifndef SDDDST_CORE_SIMULATION_H
#define SDDDST_CORE_SIMULATION_H

#include <memory>
#include <vector>

#include "dislocation.h"
#include "precision_handler.h"
#include "simulation_data.h"
#include "StressProtocols/stress_protocol.h"
#include "../sdddst_amd.h"

struct sdddst_ctx;

namespace sdddstCore {

class UmfpackLibrary;

class Simulation {
public:
    Simulation(std::shared_ptr<SimulationData> _sD);
    ~Simulation();

    void integrate(const double &stepsize, std::vector<Dislocation> &newDislocation,
                   const std::vector<Dislocation> &old, bool useSpeed2, bool calculateInitSpeed,
                   StressProtocolStepType origin, StressProtocolStepType end);
    void calculateSpeeds(const std::vector<Dislocation> &dis, std::vector<double> &res, bool ignorePHUpdate = false);
    void calculateG(const double &stepsize, std::vector<Dislocation> &newDislocation,
                    const std::vector<Dislocation> &old, bool useSpeed2, bool calculateInitSpeed,
                    bool useInitSpeedForFirstStep, StressProtocolStepType origin, StressProtocolStepType end);
    void calculateJacobian(const double &stepsize, const std::vector<Dislocation> &data);
    void calculateXError();

    void calculateSparseFormForJacobian();
    void solveEQSys();

    double calculateOrderParameter(const std::vector<double> &speeds);
    double getElement(int j, int si, int ei);
    double getSimTime();

    void run();
    bool step();
    void stepStageI();
    void stepStageII();
    void stepStageIII();

    const std::vector<Dislocation> &getStoredDislocationData();

    // The device context (for callers that drive the C ABI directly).
    sdddst_ctx *deviceContext() { return ctx; }
    // Which solver the Newton correction uses after Auto resolution.
    LinearSolver activeSolver() const { return solver; }
    const PrecisionHandler &precisionHandler() const { return *pH; }
    double getEnergy() const { return energy; }
    // Wall time spent per successful step in the last run() (seconds).
    double lastRunWallTime() const { return lastRunWall; }
    // Attempts rejected because a Newton solve did not converge and N is
    // above the dense direct solver's limit (no counterpart in the
    // reference, whose UMFPACK always solves).
    long long abortedAttemptCount() const { return abortedAttempts; }
    // Carries settings the caller changed in the shared SimulationData
    // (cutoff multiplier, precision, point-defect constants, iteration
    // count, field mode) to the device before the next step.
    void applyDataChanges();

private:
    void rejectAbortedAttempt();
    void integrateSlots(double stepsize, int newSlot, int oldSlot, bool useSpeed2, bool calculateInitSpeed,
                        StressProtocolStepType origin, StressProtocolStepType end, bool samePositions = false);
    void calculateGSlots(double stepsize, int newSlot, int oldSlot, bool useSpeed2, bool calculateInitSpeed,
                         bool useInitSpeedForFirstStep, StressProtocolStepType origin, StressProtocolStepType end);
    void speedsSlot(int slot, int buf, bool updateTolerance);
    void setStress(double t, StressProtocolStepType type, double strain);
    bool shortcutFirstIteration() const;
    void syncPositionsToHost(int slot, std::vector<Dislocation> &out);
    void uploadPositions(int slot, const std::vector<Dislocation> &in);
    void pushParams();
    void loadSpaiWeights();
    void check(int status, const char *what);
    void writeLogLine(const std::string &line);

    bool succesfulStep;
    double lastWriteTimeFinished;
    bool initSpeedCalculationIsNeeded;
    bool firstStepRequest;
    double energy;
    double energyAccum;
    double vsquare;
    double vsquare1;
    double vsquare2;
    double lastRunWall;
    long long abortedAttempts = 0;
    sdddst_params pushed{};   // what the device context holds

    std::shared_ptr<SimulationData> sD;
    std::unique_ptr<PrecisionHandler> pH;
    sdddst_ctx *ctx;
    LinearSolver solver;
    std::unique_ptr<UmfpackLibrary> umf;
    long long hostNnz;
};

}  // namespace sdddstCore

#endif
