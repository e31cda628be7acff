// sdddstCore::ProjectParser: the command line of the sdddst executable
// (project_parser.h / project_parser.cpp:29-299 of the reference), with the
// same option names, defaults, checks and exit codes, parsed without
// Boost.program_options.  Options added for the device path:
// --constant-external-stress, --linear-solver, --preconditioner,
// --field-mode, --solver-tolerance, --spai-weights.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "StressProtocols/stress_protocol.h"
#include "data_time_series.h"
#include "simulation_data.h"

namespace sdddstCore {

enum ProgramType { NONE, SIMULATION, EV_ANALYZATION };

class ProjectParser {
public:
    // Parses argv; like the reference it prints the licence (unless
    // --hide-copyright / -c), prints the help and exits 0 on --help, and
    // exits with the reference's status on an invalid command line.
    ProjectParser(int argc, char **argv);
    ~ProjectParser();

    std::shared_ptr<SimulationData> getSimulationData();
    // The EV analysis: its snapshot series (--file-list, --pd-configuration;
    // null when either is missing) and its stress protocol (a new
    // FixedRateProtocol; like the reference it aborts without
    // --fixed-rate-external-stress).
    std::shared_ptr<DataTimeSeries> getDataTimeSeries();
    StressProtocol *getStressProtocol();
    ProgramType getPType() const { return pType; }
    static void printLicense();
    static std::string helpText();

    // Parsed values (option name -> value text; flags map to "").
    const std::map<std::string, std::string> &values() const { return vm; }

private:
    void processInput();
    bool has(const std::string &name) const { return vm.count(name) != 0; }
    double asDouble(const std::string &name) const;
    unsigned asUnsigned(const std::string &name) const;
    int asInt(const std::string &name) const;
    const std::string &asString(const std::string &name) const { return vm.at(name); }

    std::map<std::string, std::string> vm;
    std::shared_ptr<SimulationData> sD;
    ProgramType pType;
};

}  // namespace sdddstCore
