// data_time_series.cpp of the reference (DataTimeSeries): the list file and
// the per-snapshot loading, with A and KASQR scaled by the snapshot's size.
#include "../../include/sdddst/data_time_series.h"

#include <cmath>
#include <fstream>
#include <sstream>

#include "../../include/sdddst/constants.h"

// The list holds one "time path" pair per line; blank lines are skipped.
DataTimeSeries::DataTimeSeries(std::string dconfList, std::string fconf) : fconfPath(fconf)
{
    std::ifstream list(dconfList);
    for (std::string line; std::getline(list, line);) {
        std::istringstream fields(line);
        std::string time, path;
        if (!(fields >> time)) continue;
        fields >> path;
        dconfs.emplace_back(std::stod(time), path);
    }
}

bool DataTimeSeries::next(std::shared_ptr<sdddstCore::SimulationData> sD)
{
    if (dconfs.empty()) return false;
    const auto pair = dconfs.front();
    sD->readDislocationDataFromFile(pair.second);
    sD->readPointDefectDataFromFile(fconfPath);
    sD->A = DEFAULT_A * 1. / sqrt(sD->dc);
    sD->KASQR = DEFAULT_KASQR * double(sD->dc);
    sD->simTime = pair.first;
    dconfs.pop_front();
    return true;
}
