// Command line of the sdddst executable: the options, defaults, checks and
// exit codes of the reference's ProjectParser (project_parser.cpp:29-299),
// parsed by a small table-driven reader (--name value, --name=value, flags).
#include "../../include/sdddst/project_parser.h"

#include <math.h>
#include <stdlib.h>

#include <iostream>
#include <sstream>

#include "../../include/sdddst/Fields/AnalyticField.h"
#include "../../include/sdddst/StressProtocols/constant_stress_protocol.h"
#include "../../include/sdddst/StressProtocols/fixed_rate_protocol.h"
#include "../../include/sdddst/StressProtocols/spring_protocol.h"
#include "../../include/sdddst_amd.h"

namespace sdddstCore {

namespace {

enum Kind { FLAG, STR, DBL, UINT, INT };

struct Spec {
    const char *name;
    Kind kind;
    const char *def;      // default value text, or nullptr
    const char *group;
    const char *help;
};

// project_parser.cpp:40-94, plus the device options at the end.
const Spec kSpecs[] = {
    {"simulation", FLAG, nullptr, "Operation", "run a simulation (default)"},
    {"ev-analyzation", FLAG, nullptr, "Operation", "run eigen value analysation"},
    {"dislocation-configuration", STR, nullptr, "Required Options",
     "plain text file path containing dislocation data in {x y b} triplets"},
    {"result-dislocation-configuration", STR, nullptr, "Required Options",
     "path where the result configuration will be stored at the end of the simulation"},
    {"point-defect-configuration", STR, nullptr, "Optional Options",
     "plain text file path containing point defect data in {x y} pairs"},
    {"logfile-path", STR, nullptr, "Optional Options",
     "path for the plain text log file (it will be overwritten if it already exists)"},
    {"time-limit", DBL, nullptr, "Optional Options", "in simulation time limit, if reached the simulation stops"},
    {"speed-limit", DBL, nullptr, "Optional Options", "in simulation units, if |v| falls below, the simulation stops"},
    {"step-count-limit", UINT, nullptr, "Optional Options", "the simulation will stop after successful N steps"},
    {"strain-increase-limit", DBL, nullptr, "Optional Options",
     "the simulation will stop after arg total strain increase is reached"},
    {"stress-limit", DBL, nullptr, "Optional Options",
     "the simulation will stop after arg external stress is reached"},
    {"avalanche-detection-limit", UINT, nullptr, "Optional Options",
     "the simulation will stop after the threshold was reached with the given numer of events from above"},
    {"avalanche-speed-threshold", DBL, "1e-3", "Optional Options", "speed threshold for counting avalanches"},
    {"initial-stepsize", DBL, "1e-6", "Optional Options", "first tried step size for the simulation"},
    {"cutoff-multiplier", DBL, "1e20", "Optional Options", "multiplier of the 1/sqrt(N) cutoff parameter"},
    {"max-stepsize", DBL, nullptr, "Optional Options", "the stepsize can not overleap this value"},
    {"calculate-strain", FLAG, nullptr, "Optional Options", "turns on strain calculation for the simulation"},
    {"calculate-order-parameter", FLAG, nullptr, "Optional Options",
     "turns on order parameter calculation during the simulation"},
    {"position-precision", DBL, "1e-5", "Optional Options",
     "minimum precision for the positions for the adaptive step size protocol"},
    {"save-sub-configurations", STR, nullptr, "Optional Options",
     "saves the current configuration after every N successful step to the given destination"},
    {"sub-configuration-delay", UINT, "5", "Optional Options",
     "number of successful steps between the sub configurations written out"},
    {"sub-configuration-delay-during-avalanche", UINT, "1", "Optional Options",
     "number of successful steps between the sub configurations written out during avalanche if avalanche detection "
     "is on"},
    {"change-cutoff-to-inf-under-threshold", DBL, nullptr, "Optional Options",
     "if the avg speed decreases once under this threshold during the simulation the applied cutoff multiplier will "
     "be 1e20"},
    {"post-relax", UINT, "0", "Optional Options", "Number of extra steps after finish condition is reached"},
    {"periodic-stress-field-analytic", FLAG, nullptr, "Available dislocation fields",
     "analytic periodic stress field (default), number of images in each direction: 4"},
    {"no-external-stress", FLAG, nullptr, "External stress protocols (optional)",
     "no external stress during the simulation (default)"},
    {"fixed-rate-external-stress", DBL, nullptr, "External stress protocols (optional)",
     "external stress is linear with time, rate should be specified as an arg"},
    {"spring-constant", DBL, nullptr, "External stress protocols (optional)",
     "simple model of an experiment where a spring is used, the arg should be the spring constant (it is valid only "
     "with the previous option together)"},
    {"constant-external-stress", DBL, nullptr, "External stress protocols (optional)",
     "constant external stress of the given value"},
    {"file-list", STR, nullptr, "EV analysis related options", "file list with timestamps int timestamp file_path format"},
    {"pd-configuration", STR, nullptr, "EV analysis related options", "file path to the point defect config file"},
    {"result-ev-file", STR, nullptr, "EV analysis related options", "file path where to save result"},
    {"calculate-ev-deriv", FLAG, nullptr, "EV analysis related options",
     "use to turn on EV derivative calculation (analytic)"},
    {"eigenvector-to-write", INT, "10", "EV analysis related options", "number of eigenvectors to write into file"},
    {"write-correlation-matrices", INT, nullptr, "EV analysis related options",
     "Write correl matrices and EV, PN into the result file only, with the given sampling size"},
    {"linear-solver", STR, "auto", "Device options",
     "auto (UMFPACK when it loads, else dense LU) | umfpack | dense-lu | fgmres"},
    {"preconditioner", STR, "pair", "Device options", "FGMRES preconditioner: pair | jacobi | learned | strong"},
    {"field-mode", STR, "fast", "Device options",
     "fast (production) | exact (reference pair terms) | ordered (reference summation order too)"},
    {"solver-tolerance", DBL, "1e-10", "Device options", "FGMRES relative residual"},
    {"spai-weights", STR, nullptr, "Device options", "weights of the learned preconditioner (raw float32)"},
    {"help", FLAG, nullptr, "Extra options", "show this help"},
    {"hide-copyright", FLAG, nullptr, "Extra options", "hides the copyright notice from the standard output"},
};

const Spec *find(const std::string &name)
{
    for (const Spec &s : kSpecs)
        if (name == s.name) return &s;
    return nullptr;
}

[[noreturn]] void fail(const std::string &msg, int code)
{
    std::cerr << msg << std::endl;
    exit(code);
}

}  // namespace

ProjectParser::ProjectParser(int argc, char **argv) : sD(nullptr), pType(NONE)
{
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        std::string name, value;
        bool inline_value = false;
        if (a == "-c") {
            name = "hide-copyright";
        } else if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                inline_value = true;
            }
        } else {
            fail("unrecognised option '" + a + "'", -1);
        }
        const Spec *s = find(name);
        if (!s) fail("unrecognised option '--" + name + "'", -1);
        if (vm.count(name)) fail("option '--" + name + "' cannot be specified more than once", -1);
        if (s->kind == FLAG) {
            if (inline_value) fail("option '--" + name + "' does not take any arguments", -1);
            vm[name] = "";
            continue;
        }
        if (!inline_value) {
            if (i + 1 >= argc) fail("the required argument for option '--" + name + "' is missing", -1);
            value = argv[++i];
        }
        vm[name] = value;
        // validate numbers now, as the reference's parser does
        if (s->kind == DBL) (void)asDouble(name);
        if (s->kind == UINT) (void)asUnsigned(name);
        if (s->kind == INT) (void)asInt(name);
    }
    for (const Spec &s : kSpecs)
        if (s.def && !vm.count(s.name)) vm[s.name] = s.def;

    if (!has("hide-copyright")) printLicense();
    if (has("help")) {
        std::cout << helpText() << std::endl;
        exit(0);
    }
    processInput();
}

ProjectParser::~ProjectParser() = default;

std::shared_ptr<SimulationData> ProjectParser::getSimulationData() { return sD; }

void ProjectParser::printLicense()
{
    std::cout << "SDDDST - Simple Discrete Dislocation Dynamics Toolkit (MI355X device path)\n"
                 "This program comes with ABSOLUTELY NO WARRANTY; This is free software, and you are welcome to "
                 "redistribute it under certain conditions; see the license for details\n";
}

std::string ProjectParser::helpText()
{
    std::ostringstream o;
    std::string group;
    for (const Spec &s : kSpecs) {
        if (group != s.group) {
            group = s.group;
            o << "\n" << group << ":\n";
        }
        std::string left = std::string("  --") + s.name;
        if (s.kind != FLAG) left += " arg";
        if (s.def) left += std::string(" (=") + s.def + ")";
        o << left;
        if (left.size() < 48) o << std::string(48 - left.size(), ' ');
        else o << "\n" << std::string(48, ' ');
        o << s.help << "\n";
    }
    return o.str();
}

double ProjectParser::asDouble(const std::string &name) const
{
    const std::string &v = vm.at(name);
    char *end = nullptr;
    const double d = strtod(v.c_str(), &end);
    if (v.empty() || *end != '\0') fail("the argument ('" + v + "') for option '--" + name + "' is invalid", -1);
    return d;
}

unsigned ProjectParser::asUnsigned(const std::string &name) const
{
    const std::string &v = vm.at(name);
    char *end = nullptr;
    const unsigned long u = strtoul(v.c_str(), &end, 10);
    if (v.empty() || *end != '\0' || v[0] == '-' || u > 0xffffffffUL)
        fail("the argument ('" + v + "') for option '--" + name + "' is invalid", -1);
    return (unsigned)u;
}

int ProjectParser::asInt(const std::string &name) const
{
    const std::string &v = vm.at(name);
    char *end = nullptr;
    const long l = strtol(v.c_str(), &end, 10);
    if (v.empty() || *end != '\0') fail("the argument ('" + v + "') for option '--" + name + "' is invalid", -1);
    return (int)l;
}

// project_parser.cpp:137-299
void ProjectParser::processInput()
{
    if (!has("ev-analyzation")) {
        pType = SIMULATION;
        if (!has("dislocation-configuration")) fail("dislocation-configuration is missing!", -1);
        if (!has("result-dislocation-configuration")) fail("result-dislocation-configuration is missing!", -1);
        sD = std::make_shared<SimulationData>(asString("dislocation-configuration"),
                                              has("point-defect-configuration")
                                                  ? asString("point-defect-configuration")
                                                  : std::string());
        if (has("time-limit")) {
            sD->isTimeLimit = true;
            sD->timeLimit = asDouble("time-limit");
        }
        sD->stepSize = asDouble("initial-stepsize");
        if (has("stress-limit")) {
            sD->stressLimit = asDouble("stress-limit");
            sD->isStressLimit = true;
        }
        sD->prec = asDouble("position-precision");
        sD->cutOffMultiplier = asDouble("cutoff-multiplier");
        sD->updateCutOff();
        if (has("strain-increase-limit")) {
            sD->isStrainIncreaseLimit = true;
            sD->totalAccumulatedStrainIncreaseLimit = asDouble("strain-increase-limit");
        }
        if (has("max-stepsize")) {
            sD->maxStepSizeLimit = asDouble("max-stepsize");
            sD->isMaxStepSizeLimit = true;
        }
        if (has("point-defect-configuration")) {
            sD->A *= 1. / sqrt(sD->dc);
            sD->KASQR *= double(sD->dc);
        }
        if (has("calculate-strain")) sD->calculateStrainDuringSimulation = true;
        if (has("calculate-order-parameter")) sD->orderParameterCalculationIsOn = true;
        if (has("logfile-path")) sD->standardOutputLog = std::ofstream(asString("logfile-path"));
        if (has("step-count-limit")) {
            sD->isStepCountLimit = true;
            sD->stepCountLimit = asUnsigned("step-count-limit");
        }
        if (has("avalanche-detection-limit")) {
            sD->countAvalanches = true;
            sD->avalancheTriggerLimit = asUnsigned("avalanche-detection-limit");
            sD->avalancheSpeedThreshold = asDouble("avalanche-speed-threshold");
        }
        if (has("save-sub-configurations")) {
            sD->isSaveSubConfigs = true;
            sD->subConfigPath = asString("save-sub-configurations");
            sD->subConfigDelay = asUnsigned("sub-configuration-delay");
            sD->subConfigDelayDuringAvalanche = asUnsigned("sub-configuration-delay-during-avalanche");
        }
        if (has("speed-limit")) {
            sD->isSpeedLimit = true;
            sD->speedLimit = asDouble("speed-limit");
        }
        sD->remainingFinalSteps = asUnsigned("post-relax");
        sD->endDislocationConfigurationPath = asString("result-dislocation-configuration");
        sD->tau = std::unique_ptr<Field>(new AnalyticField());

        const int protocols = (int)has("no-external-stress") + (int)has("fixed-rate-external-stress") +
                              (int)has("constant-external-stress");
        if (protocols > 1) fail("Only one stress protocol can be defined at the same time!", -20);
        if (has("fixed-rate-external-stress")) {
            std::unique_ptr<FixedRateProtocol> tmp;
            if (has("spring-constant")) {
                std::unique_ptr<SpringProtocol> spring(new SpringProtocol);
                spring->setSpringConstant(asDouble("spring-constant"));
                tmp.reset(spring.release());
            } else {
                tmp.reset(new FixedRateProtocol);
            }
            tmp->setRate(asDouble("fixed-rate-external-stress"));
            sD->externalStressProtocol = std::unique_ptr<StressProtocol>(std::move(tmp));
        } else if (has("constant-external-stress")) {
            sD->externalStressProtocol.reset(new ConstantStressProtocol(asDouble("constant-external-stress")));
        } else {
            sD->externalStressProtocol = std::unique_ptr<StressProtocol>(new StressProtocol());
        }
        if (has("change-cutoff-to-inf-under-threshold")) {
            sD->speedThresholdForCutoffChange = asDouble("change-cutoff-to-inf-under-threshold");
            sD->isSpeedThresholdForCutoffChange = true;
        }

        // device options
        const std::string ls = asString("linear-solver");
        if (ls == "auto") sD->linearSolver = LinearSolver::Auto;
        else if (ls == "umfpack") sD->linearSolver = LinearSolver::Umfpack;
        else if (ls == "dense-lu") sD->linearSolver = LinearSolver::DenseLU;
        else if (ls == "fgmres") sD->linearSolver = LinearSolver::FGMRES;
        else fail("the argument ('" + ls + "') for option '--linear-solver' is invalid", -1);
        const std::string pc = asString("preconditioner");
        if (pc == "pair") sD->preconditioner = SDDDST_PRECOND_PAIR;
        else if (pc == "jacobi") sD->preconditioner = SDDDST_PRECOND_JACOBI;
        else if (pc == "learned") sD->preconditioner = SDDDST_PRECOND_LEARNED;
        else if (pc == "strong") sD->preconditioner = SDDDST_PRECOND_STRONG;
        else fail("the argument ('" + pc + "') for option '--preconditioner' is invalid", -1);
        const std::string fm = asString("field-mode");
        if (fm == "fast") sD->fieldMode = SDDDST_FIELD_FAST;
        else if (fm == "exact") sD->fieldMode = SDDDST_FIELD_EXACT;
        else if (fm == "ordered") sD->fieldMode = SDDDST_FIELD_ORDERED;
        else fail("the argument ('" + fm + "') for option '--field-mode' is invalid", -1);
        sD->solverRelTolerance = asDouble("solver-tolerance");
        if (has("spai-weights")) sD->spaiWeightsPath = asString("spai-weights");
    } else if (!has("simulation")) {
        // project_parser.cpp:298-319
        sD = std::make_shared<SimulationData>();
        if (!has("result-ev-file")) fail("result-ev-file is not defined! Exiting...", 1);
        sD->eVAnalResultFile = asString("result-ev-file");
        pType = EV_ANALYZATION;
        if (has("calculate-ev-deriv")) sD->calculateDerivativeEVAnal = true;
        if (has("eigenvector-to-write")) sD->numberOfEigenVecToWrite = asInt("eigenvector-to-write");
        if (has("write-correlation-matrices")) sD->writeCorrelMatrices = asInt("write-correlation-matrices");
    }
}

// project_parser.cpp:330-338
std::shared_ptr<DataTimeSeries> ProjectParser::getDataTimeSeries()
{
    if (pType == EV_ANALYZATION && has("pd-configuration") && has("file-list"))
        return std::make_shared<DataTimeSeries>(asString("file-list"), asString("pd-configuration"));
    return nullptr;
}

// project_parser.cpp:340-348: the analysis takes a fixed-rate protocol only
StressProtocol *ProjectParser::getStressProtocol()
{
    if (has("fixed-rate-external-stress")) {
        auto *fpr = new FixedRateProtocol();
        fpr->setRate(asDouble("fixed-rate-external-stress"));
        return fpr;
    }
    abort();
}

}  // namespace sdddstCore
