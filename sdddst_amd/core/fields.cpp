// Field / AnalyticField (reference: src/Fields/Field.cpp:22-45,
// src/Fields/AnalyticField.cpp:134-202).  AnalyticField's per-point methods
// evaluate the device implementation (the same code the kernels run) through
// a small private context; Simulation never calls them per pair.
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/sdddst/Fields/AnalyticField.h"
#include "../../include/sdddst/Fields/Field.h"
#include "../../include/sdddst_amd.h"

using namespace sdddstCore;

Field::Field() {}
Field::~Field() {}
double Field::xy(double, double) { return 0; }
double Field::xy_diff_x(double, double) { return 0; }
double Field::xy_diff_x2(double, double) { return 0; }

namespace {
sdddst_ctx *point_context()
{
    static std::mutex mu;
    static sdddst_ctx *ctx = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx) {
        sdddst_params p{1.0, 1e-6, 0.0, 0.0, 2, SDDDST_FIELD_EXACT};
        if (sdddst_ctx_create(2, 0, &p, &ctx) != SDDDST_OK)
            throw std::runtime_error(std::string("AnalyticField: ") + sdddst_last_error());
    }
    return ctx;
}

double eval_point(int which, double dx, double dy)
{
    double out = 0;
    if (sdddst_field_eval(point_context(), which, 1, &dx, &dy, &out) != SDDDST_OK)
        throw std::runtime_error(std::string("AnalyticField: ") + sdddst_last_error());
    return out;
}
}  // namespace

AnalyticField::AnalyticField() : Field() {}
AnalyticField::~AnalyticField() {}
double AnalyticField::xy(double dx, double dy) { return eval_point(0, dx, dy); }
double AnalyticField::xy_diff_x(double dx, double dy) { return eval_point(1, dx, dy); }
double AnalyticField::xy_diff_x2(double dx, double dy) { return eval_point(3, dx, dy); }
