// Adaptive precision / step-size controller (reference:
// include/precision_handler.h:28-60, src/precision_handler.cpp).  The
// per-dislocation tolerance and error live on the device and are updated by
// the kernels; this object holds the scalar state the host sees.
#ifndef SDDDST_CORE_PRECISION_HANDLER_H
#define SDDDST_CORE_PRECISION_HANDLER_H

#include <string>
#include <utility>
#include <vector>

namespace sdddstCore {

class PrecisionHandler {
public:
    PrecisionHandler();
    ~PrecisionHandler();

    void setSize(unsigned int size);
    unsigned long getSize() const;

    // precision_handler.cpp:62-70 (the device resets its own arrays)
    void reset();
    // precision_handler.cpp:72-101 on host arrays of size getSize(), for
    // callers of the class API; the simulator updates its tolerance and
    // error on the device.
    void updateTolerance(const double &distanceSqr, const unsigned int &ID);
    void updateError(const double &error, const unsigned int &ID);
    // precision_handler.cpp:103-119, with the host C library's pow.
    double getNewStepSize(const double &oldStepSize) const;

    double getMinPrecisity() const;
    void setMinPrecisity(double value);

    double getMaxErrorRatioSqr() const;
    unsigned int getSelectedID() const;

    // Set by the device error controller.
    void setFromDevice(double maxErrorRatioSqr, unsigned int selectedID);

    std::string __str__() const;

private:
    unsigned long size;
    double minPrecisity;
    double minPrecisitySqr;
    double maxErrorRatioSqr;
    unsigned int selectedID;
    std::vector<std::pair<double, double>> toleranceAndError;
    bool hostTracking = false;
};

}  // namespace sdddstCore

#endif
