// SimulationData (reference: src/simulation_data.cpp:43-298).
#include <math.h>
#include <stdlib.h>

#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "../../include/sdddst/constants.h"
#include "../../include/sdddst/simulation_data.h"
#include "../../include/sdddst_amd.h"

using namespace sdddstCore;

SimulationData::SimulationData()
    : cutOffMultiplier(DEFAULT_CUTOFF_MULTIPLIER), cutOff(DEFAULT_CUTOFF), cutOffSqr(cutOff * cutOff),
      onePerCutOffSqr(1. / cutOffSqr), prec(DEFAULT_PRECISION), pc(0), dc(0), ic(DEFAULT_ITERATION_COUNT),
      timeLimit(DEFAULT_TIME_LIMIT), stepSize(DEFAULT_STEP_SIZE), simTime(DEFAULT_SIM_TIME), KASQR(DEFAULT_KASQR),
      A(DEFAULT_A), tau(nullptr), Ap(nullptr), Ai(nullptr), Ax(nullptr), x(nullptr), null(nullptr),
      Symbolic(nullptr), Numeric(nullptr), succesfulSteps(0), failedSteps(0), totalAccumulatedStrainIncrease(0),
      isStrainIncreaseLimit(false), totalAccumulatedStrainIncreaseLimit(0), isMaxStepSizeLimit(false),
      maxStepSizeLimit(0), isTimeLimit(false), isStepCountLimit(false), stepCountLimit(0), isSpeedLimit(false),
      finish(false), speedLimit(0), stressLimit(0), isStressLimit(false), calculateStrainDuringSimulation(false),
      orderParameterCalculationIsOn(false), standardOutputLog(), endDislocationConfigurationPath(""),
      externalStressProtocol(nullptr), countAvalanches(false), avalancheSpeedThreshold(0), avalancheTriggerLimit(0),
      avalancheCount(0), inAvalanche(false), isSaveSubConfigs(false), subConfigPath(""), subConfigDelay(0),
      subConfigDelayDuringAvalanche(0), subconfigDistanceCounter(0), currentStressStateType(Original),
      speedThresholdForCutoffChange(0), isSpeedThresholdForCutoffChange(false), currentStorageSize(0),
      sumAvgSpeed(0), eVAnalResultFile(""), remainingFinalSteps(0), inFinal(false), calculateDerivativeEVAnal(false),
      numberOfEigenVecToWrite(10), writeCorrelMatrices(0), linearSolver(LinearSolver::Auto),
      solverRelTolerance(1e-10), solverMaxIterations(300), solverRestart(40), preconditioner(SDDDST_PRECOND_PAIR),
      fieldMode(SDDDST_FIELD_FAST),
      dislocationDataIsLoaded(false)
{
}

SimulationData::SimulationData(const std::string &dislocationDataFilePath, const std::string &pointDefectDataFilePath)
    : SimulationData()
{
    readDislocationDataFromFile(dislocationDataFilePath);
    readPointDefectDataFromFile(pointDefectDataFilePath);
    initSimulationVariables();
}

SimulationData::~SimulationData() { deleteDislocationCountRelatedData(); }

namespace {
// The configuration files are plain whitespace-separated numbers, `width`
// per record ({x y b} for dislocations, {x y} for point defects; the
// reference reads them at simulation_data.cpp:119-150, 163-188).  Every
// token is converted as a decimal double; a token that is not one, or a
// count that is not a whole number of records, is an error naming the file.
std::vector<double> readRecords(const std::string &path, size_t width, const char *what)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error(std::string("Cannot open ") + what + " file: " + path);
    std::vector<double> values;
    for (std::string token; in >> token;) {
        size_t used = 0;
        double v = 0;
        try {
            v = std::stod(token, &used);
        } catch (const std::exception &) {
            used = 0;
        }
        if (used != token.size())
            throw std::runtime_error(std::string(what) + " file " + path + ": '" + token + "' is not a number");
        values.push_back(v);
    }
    if (values.size() % width != 0)
        throw std::runtime_error(std::string(what) + " file " + path + ": " + std::to_string(values.size()) +
                                 " numbers are not whole records of " + std::to_string(width));
    return values;
}
}  // namespace

void SimulationData::readDislocationDataFromFile(const std::string &path)
{
    if (path.empty()) return;
    const std::vector<double> v = readRecords(path, 3, "dislocation data");
    std::vector<Dislocation> d(v.size() / 3);
    for (size_t i = 0; i < d.size(); ++i) d[i] = Dislocation{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    setDislocations(d);
}

// simulation_data.cpp:152-161
void SimulationData::writeDislocationDataToFile(const std::string &path)
{
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("Cannot open the data file to write: " + path);
    out << std::scientific << std::setprecision(16);
    for (auto &i : dislocations) out << i.x << " " << i.y << " " << i.b << "\n";
}

void SimulationData::readPointDefectDataFromFile(const std::string &path)
{
    pc = 0;
    points.clear();
    if (path.empty()) return;
    const std::vector<double> v = readRecords(path, 2, "fixpoints data");
    std::vector<PointDefect> p(v.size() / 2);
    for (size_t i = 0; i < p.size(); ++i) p[i] = PointDefect{v[2 * i], v[2 * i + 1]};
    setPointDefects(p);
}

// simulation_data.cpp:190-201
void SimulationData::writePointDefectDataToFile(const std::string &path)
{
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("Cannot open the data file to write: " + path);
    out << std::scientific << std::setprecision(16);
    for (auto &i : points) out << i.x << " " << i.y << "\n";
}

void SimulationData::setDislocations(const std::vector<Dislocation> &d)
{
    deleteDislocationCountRelatedData();
    dislocations = d;
    dc = (unsigned int)d.size();
    dislocationDataIsLoaded = true;
    updateMemoryUsageAccordingToDislocationCount();
}

void SimulationData::setPointDefects(const std::vector<PointDefect> &p)
{
    points = p;
    pc = (unsigned int)p.size();
}

void SimulationData::initSimulationVariables() { updateCutOff(); }

// simulation_data.cpp:208-218
void SimulationData::updateCutOff()
{
    double multiplier = cutOffMultiplier;
    if (cutOffMultiplier >= 1. / 12. * sqrt(2.0 * double(dc))) multiplier = 1e20;
    cutOff = 1. / sqrt(dc) * multiplier;
    cutOffSqr = cutOff * cutOff;
    onePerCutOffSqr = 1. / cutOffSqr;
}

void SimulationData::deleteDislocationCountRelatedData()
{
    g.clear();
    initSpeed.clear();
    initSpeed2.clear();
    speed.clear();
    speed2.clear();
    dVec.clear();
    bigStep.clear();
    firstSmall.clear();
    secondSmall.clear();
    free(Ap);
    Ap = nullptr;
    free(Ai);
    Ai = nullptr;
    free(Ax);
    Ax = nullptr;
    free(x);
    x = nullptr;
    indexes.clear();
    currentStorageSize = 0;
}

// simulation_data.cpp:277-298; the CSC index/value arrays are sized to nnz
// when the Jacobian is downloaded, not N^2.
void SimulationData::updateMemoryUsageAccordingToDislocationCount()
{
    g.assign(dc, 0.0);
    initSpeed.assign(dc, 0.0);
    initSpeed2.assign(dc, 0.0);
    speed.assign(dc, 0.0);
    speed2.assign(dc, 0.0);
    dVec.assign(dc, 0.0);
    bigStep = dislocations;
    firstSmall = dislocations;
    secondSmall = dislocations;
    Ap = (int *)calloc(dc + 1, sizeof(int));
    x = (double *)calloc(dc ? dc : 1, sizeof(double));
    if (!Ap || !x) throw std::bad_alloc();
    indexes.assign(dc, 0);
}
