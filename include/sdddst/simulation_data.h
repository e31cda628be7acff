// Simulation state container (reference: include/simulation_data.h:38-281,
// src/simulation_data.cpp).  Same public fields; the device holds the live
// copies of positions, speeds and the Jacobian while Simulation runs, and
// the host fields below are refreshed at the API boundary
// (getStoredDislocationData, calculateJacobian, the log, file output).
#ifndef SDDDST_CORE_SIMULATION_DATA_H
#define SDDDST_CORE_SIMULATION_DATA_H

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "Fields/Field.h"
#include "StressProtocols/stress_protocol.h"
#include "dislocation.h"

namespace sdddstCore {

// How the Newton correction is solved (new in this build).
enum class LinearSolver {
    Auto,       // UMFPACK on the host when libumfpack loads, else DenseLU
    Umfpack,    // host UMFPACK hand-off exactly as the reference
    DenseLU,    // device LU with partial pivoting
    FGMRES      // device FGMRES (relative residual tolerance), DenseLU fallback
};

class SimulationData {
public:
    SimulationData(const std::string &dislocationDataFilePath, const std::string &pointDefectDataFilePath);
    SimulationData();
    ~SimulationData();

    // simulation_data.cpp:119-201
    void readDislocationDataFromFile(const std::string &dislocationDataFilePath);
    void writeDislocationDataToFile(const std::string &dislocationDataFilePath);
    void readPointDefectDataFromFile(const std::string &pointDefectDataFilePath);
    void writePointDefectDataToFile(const std::string &pointDefectDataFilePath);

    // Direct construction from arrays (positions, Burgers vectors, defects).
    void setDislocations(const std::vector<Dislocation> &d);
    void setPointDefects(const std::vector<PointDefect> &p);

    void initSimulationVariables();
    void updateCutOff();   // simulation_data.cpp:208-218

    std::vector<Dislocation> dislocations;
    std::vector<PointDefect> points;
    std::vector<double> g;
    std::vector<double> initSpeed;
    std::vector<double> initSpeed2;
    std::vector<double> speed;
    std::vector<double> speed2;
    std::vector<double> dVec;
    double cutOffMultiplier;
    double cutOff;
    double cutOffSqr;
    double onePerCutOffSqr;
    double prec;
    unsigned int pc;
    unsigned int dc;
    unsigned int ic;
    double timeLimit;
    double stepSize;
    double simTime;
    double KASQR;
    double A;
    std::vector<Dislocation> bigStep;
    std::vector<Dislocation> firstSmall;
    std::vector<Dislocation> secondSmall;
    std::unique_ptr<Field> tau;
    // CSC of the Jacobian as UMFPACK receives it (host copies)
    int *Ap;
    int *Ai;
    double *Ax;
    double *x;
    double *null;
    void *Symbolic, *Numeric;
    std::vector<int> indexes;
    size_t succesfulSteps;
    size_t failedSteps;
    double totalAccumulatedStrainIncrease;
    bool isStrainIncreaseLimit;
    double totalAccumulatedStrainIncreaseLimit;
    bool isMaxStepSizeLimit;
    double maxStepSizeLimit;
    bool isTimeLimit;
    bool isStepCountLimit;
    unsigned int stepCountLimit;
    bool isSpeedLimit;
    bool finish;
    double speedLimit;
    double stressLimit;
    bool isStressLimit;
    bool calculateStrainDuringSimulation;
    bool orderParameterCalculationIsOn;
    std::ofstream standardOutputLog;
    std::string endDislocationConfigurationPath;
    std::unique_ptr<StressProtocol> externalStressProtocol;
    bool countAvalanches;
    double avalancheSpeedThreshold;
    unsigned int avalancheTriggerLimit;
    unsigned int avalancheCount;
    bool inAvalanche;
    bool isSaveSubConfigs;
    std::string subConfigPath;
    unsigned int subConfigDelay;
    unsigned int subConfigDelayDuringAvalanche;
    unsigned int subconfigDistanceCounter;
    StressProtocolStepType currentStressStateType;
    double speedThresholdForCutoffChange;
    bool isSpeedThresholdForCutoffChange;
    unsigned currentStorageSize;
    double sumAvgSpeed;
    std::string eVAnalResultFile;
    unsigned int remainingFinalSteps;
    bool inFinal;
    bool calculateDerivativeEVAnal;
    int numberOfEigenVecToWrite;
    int writeCorrelMatrices;

    // Device execution options (new in this build).
    LinearSolver linearSolver;
    double solverRelTolerance;    // FGMRES acceptance ||g - A d|| / ||g||
    int solverMaxIterations;
    int solverRestart;
    int preconditioner;           // SDDDST_PRECOND_* of include/sdddst_amd.h
    std::string spaiWeightsPath;  // learned preconditioner weights; empty: the shipped file
    int fieldMode;                // SDDDST_FIELD_* of include/sdddst_amd.h

private:
    void updateMemoryUsageAccordingToDislocationCount();
    void deleteDislocationCountRelatedData();
    bool dislocationDataIsLoaded;
};

}  // namespace sdddstCore

#endif
