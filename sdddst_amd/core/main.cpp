// The sdddst executable (main.cpp of the reference): parse the command line,
// run the simulation on the device.  The eigenvalue analysis
// (time_series_processor.cpp) is not part of the device path.
#include <exception>
#include <iostream>

#include "../../include/sdddst/project_parser.h"
#include "../../include/sdddst/simulation.h"

int main(int argc, char **argv)
{
    sdddstCore::ProjectParser parser(argc, argv);
    if (parser.getPType() == sdddstCore::SIMULATION) {
        try {
            sdddstCore::Simulation simulation(parser.getSimulationData());
            simulation.run();
        } catch (const std::exception &e) {
            std::cerr << "sdddst: " << e.what() << std::endl;
            return 1;
        }
    } else if (parser.getPType() == sdddstCore::EV_ANALYZATION) {
        std::cerr << "sdddst: the eigenvalue analysis is not available in this build" << std::endl;
        return 3;
    }
    return 0;
}
