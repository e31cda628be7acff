// Per-element and per-system formulas of one implicit step, shared by the
// single-system kernels (step.hip) and the batched ensemble (ensemble.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "libm_compat.h"

namespace sdamd {
namespace stepf {

// Simulation::calculateG (simulation.cpp:189-192)
__device__ __forceinline__ double residual(double xn, double xo, double csp, double isp, double d, double h)
{
    return xn - (1.0 + d) * 0.5 * h * csp - xo - (1.0 - d) * 0.5 * h * isp;
}

// PrecisionHandler::getNewStepSize (precision_handler.cpp:103-119) and the
// max-stepsize clamp (simulation.cpp:645-648); hmax < 0: no limit.
__device__ __forceinline__ double new_step_size(double maxRatio, double h, double hmax)
{
    double hn;
    if (0.0 == maxRatio) {
        hn = h * 2.0;
    } else {
        double tmp = 1. / sdlibm::sqrt(maxRatio);
        tmp = sdlibm::pow_one_third(tmp);
        if (tmp > 2.0) tmp = 2.0;
        hn = 0.9 * h * tmp;
    }
    if (hmax >= 0 && hmax < hn) hn = hmax;
    return hn;
}

// PrecisionHandler::updateTolerance (precision_handler.cpp:72-87)
__device__ __forceinline__ void tol_update(double &tol, double distanceSqr, double minPrecSqr)
{
    const double tmp = distanceSqr * 0.25 * 1e-2;
    if (tmp < minPrecSqr && tmp < tol) tol = (tmp == 0.0) ? 1e-12 : tmp;
}

}  // namespace stepf
}  // namespace sdamd
