// The EV analysis of a time series (time_series_processor.h of the
// reference): per snapshot the eigen-decomposition of the dense Hessian on
// the device (sdddst_ev_analysis, with the eigenvalue derivatives of
// --calculate-ev-deriv when asked for) and the reference's gzip-compressed
// text output.
#pragma once

#include <memory>

#include "StressProtocols/stress_protocol.h"
#include "data_time_series.h"
#include "simulation_data.h"

namespace sdddstEV {

class TimeSeriesProcessor {
public:
    TimeSeriesProcessor(std::shared_ptr<sdddstCore::SimulationData> simData,
                        std::shared_ptr<DataTimeSeries> dateTimeSeries, sdddstCore::StressProtocol *stressProtocol);
    void run();

private:
    std::shared_ptr<sdddstCore::SimulationData> sD;
    std::shared_ptr<DataTimeSeries> dTS;
};

}  // namespace sdddstEV
