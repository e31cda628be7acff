// Host helpers of the simulator API (reference: include/utility.h:29-100).
// The device kernels carry their own copies of the arithmetic; these are the
// host-side functions of the public API.
#ifndef SDDDST_CORE_UTILITY_H
#define SDDDST_CORE_UTILITY_H

#include <math.h>
#include <sys/time.h>
#include <time.h>

#include <vector>

#include "dislocation.h"

// Wrap into [-0.5, 0.5) (utility.h:29-40).
inline void normalize(double &n)
{
    while (n < -0.5) n += 1.0;
    while (n >= 0.5) n -= 1.0;
}

// Point-defect shape functions (utility.h:42-70).
inline double X(const double &x) { return sin(2.0 * M_PI * x) * 0.5 / M_PI; }
inline double X2(const double &x) { return (1.0 - cos(2.0 * M_PI * x)) * 0.5 / M_PI / M_PI; }
inline double E(const double &x, const double &y, const double &K) { return exp(-K * (X2(x) + X2(y))); }
inline double X_dx(const double &x) { return cos(2.0 * M_PI * x); }
inline double X2_dx(const double &x) { return sin(2.0 * M_PI * x) / M_PI; }
inline double E_dx(const double &x, const double &y, const double &K) { return -E(x, y, K) * K * X2_dx(x); }

// Wall and CPU clocks (utility.h:75-90).
inline double get_wall_time()
{
    struct timeval t;
    if (gettimeofday(&t, nullptr)) return 0;
    return (double)t.tv_sec + (double)t.tv_usec * .000001;
}
inline double get_cpu_time() { return (double)clock() / CLOCKS_PER_SEC; }

// Sum of b * (x_new - x_old) in index order (utility.h:92-100).
inline double calculateStrainIncrement(const std::vector<sdddstCore::Dislocation> &old,
                                       const std::vector<sdddstCore::Dislocation> &newD)
{
    double result = 0;
    for (size_t i = 0; i < old.size(); i++) result += newD[i].b * (newD[i].x - old[i].x);
    return result;
}

#endif
