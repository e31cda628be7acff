// PrecisionHandler scalar state (reference: src/precision_handler.cpp:28-151).
#include <math.h>

#include <iostream>
#include <string>

#include "../../include/sdddst/constants.h"
#include "../../include/sdddst/precision_handler.h"

using namespace sdddstCore;

PrecisionHandler::PrecisionHandler()
    : size(0), minPrecisity(1e-6), minPrecisitySqr(1e-12), maxErrorRatioSqr(0), selectedID(0)
{
}

PrecisionHandler::~PrecisionHandler() {}

void PrecisionHandler::setSize(unsigned int s)
{
    size = s;
    // entries kept, new ones at (prec^2, 0) (precision_handler.cpp:42-55)
    toleranceAndError.resize(s, std::make_pair(minPrecisitySqr, 0.0));
}
unsigned long PrecisionHandler::getSize() const { return size; }

void PrecisionHandler::reset()
{
    maxErrorRatioSqr = 0;
    // the host arrays are touched only by callers of updateTolerance /
    // updateError; the simulator's own arrays live on the device
    if (hostTracking)
        for (auto &te : toleranceAndError) te = std::make_pair(minPrecisitySqr, 0.0);
}

// precision_handler.cpp:72-87
void PrecisionHandler::updateTolerance(const double &distanceSqr, const unsigned int &ID)
{
    hostTracking = true;
    const double candidate = distanceSqr * 0.25 * 1e-2;
    double &tol = toleranceAndError.at(ID).first;
    if (candidate < minPrecisitySqr && candidate < tol) {
        if (candidate == 0.0) {
            tol = EPS;
            std::cout << "Two dislocations are in the same place!\n";
        } else {
            tol = candidate;
        }
    }
}

// precision_handler.cpp:89-101
void PrecisionHandler::updateError(const double &error, const unsigned int &ID)
{
    hostTracking = true;
    std::pair<double, double> &te = toleranceAndError.at(ID);
    if (te.second < error) {
        te.second = error;
        const double ratio = error * error / te.first;
        if (ratio > maxErrorRatioSqr) {
            maxErrorRatioSqr = ratio;
            selectedID = ID;
        }
    }
}

double PrecisionHandler::getNewStepSize(const double &oldStepSize) const
{
    if (0.0 == maxErrorRatioSqr) return oldStepSize * 2.0;
    double tmp = 1. / sqrt(maxErrorRatioSqr);
    tmp = pow(tmp, 1. / 3.);
    if (tmp > 2.0) tmp = 2.0;
    return 0.9 * oldStepSize * tmp;
}

double PrecisionHandler::getMinPrecisity() const { return minPrecisity; }

void PrecisionHandler::setMinPrecisity(double value)
{
    minPrecisity = value;
    minPrecisitySqr = value * value;
}

double PrecisionHandler::getMaxErrorRatioSqr() const { return maxErrorRatioSqr; }
unsigned int PrecisionHandler::getSelectedID() const { return selectedID; }

void PrecisionHandler::setFromDevice(double m, unsigned int id)
{
    maxErrorRatioSqr = m;
    selectedID = id;
}

std::string PrecisionHandler::__str__() const
{
    return "N: " + std::to_string(size) + " Prec: " + std::to_string(minPrecisity) +
           " New stepsize factor: " + std::to_string(getNewStepSize(1.0));
}
