// The snapshots of an EV analysis (data_time_series.h of the reference): a
// list file of "time path" pairs; each next() loads one dislocation
// configuration and the point-defect configuration into SimulationData.
#pragma once

#include <deque>
#include <memory>
#include <string>
#include <utility>

#include "simulation_data.h"

class DataTimeSeries {
public:
    DataTimeSeries(std::string dconfList, std::string fconf);
    bool next(std::shared_ptr<sdddstCore::SimulationData> sD);

private:
    std::deque<std::pair<double, std::string>> dconfs;
    std::string fconfPath;
};
