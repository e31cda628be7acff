// PrecisionHandler scalar state (reference: src/precision_handler.cpp:28-151).
#include <math.h>

#include <string>

#include "../../include/sdddst/precision_handler.h"

using namespace sdddstCore;

PrecisionHandler::PrecisionHandler()
    : size(0), minPrecisity(1e-6), minPrecisitySqr(1e-12), maxErrorRatioSqr(0), selectedID(0)
{
}

PrecisionHandler::~PrecisionHandler() {}

void PrecisionHandler::setSize(unsigned int s) { size = s; }
unsigned long PrecisionHandler::getSize() const { return size; }

void PrecisionHandler::reset() { maxErrorRatioSqr = 0; }

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
