// The sdddst executable (main.cpp of the reference): parse the command line,
// run the simulation or the eigenvalue analysis of a time series on the device.
#include <exception>
#include <iostream>

#include "../../include/sdddst/project_parser.h"
#include "../../include/sdddst/simulation.h"
#include "../../include/sdddst/time_series_processor.h"

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
        try {
            sdddstEV::TimeSeriesProcessor processor(parser.getSimulationData(), parser.getDataTimeSeries(),
                                                    parser.getStressProtocol());
            processor.run();
        } catch (const std::exception &e) {
            std::cerr << "sdddst: " << e.what() << std::endl;
            return 1;
        }
    }
    return 0;
}
