// C entry points over sdddstCore::Simulation (include/sdddst/core_capi.h).
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sdddst/Fields/AnalyticField.h"
#include "../../include/sdddst/StressProtocols/constant_stress_protocol.h"
#include "../../include/sdddst/StressProtocols/fixed_rate_protocol.h"
#include "../../include/sdddst/StressProtocols/spring_protocol.h"
#include "../../include/sdddst/core_capi.h"
#include "../../include/sdddst/simulation.h"
#include "../../include/sdddst/utility.h"
#include "../../include/sdddst_amd.h"

using namespace sdddstCore;

struct sdddst_data;

struct sdddst_sim {
    std::shared_ptr<SimulationData> sD;
    std::unique_ptr<Simulation> sim;
    bool lastSuccess = false;
    double lastWall = 0;
    sdddst_data *owner = nullptr;   // the data handle it was created from, while both live
};

// A SimulationData shared with the Simulations created from it, as the
// bindings' SimulationDataObject holds it (bindings.cpp:196-206).
struct sdddst_data {
    std::shared_ptr<SimulationData> sD;
    sdddst_sim *sim = nullptr;   // the last Simulation created from it, for live positions
};

namespace {
thread_local std::string core_error;

int fail(const std::exception &e)
{
    core_error = e.what();
    return SDDDST_ERR_STATE;
}

void configure(SimulationData &d, const sdddst_sim_options &o)
{
    d.prec = o.precision;
    d.stepSize = o.initial_step;
    d.cutOffMultiplier = o.cutoff_multiplier;
    d.updateCutOff();
    if (o.max_step >= 0) {
        d.isMaxStepSizeLimit = true;
        d.maxStepSizeLimit = o.max_step;
    }
    d.calculateStrainDuringSimulation = o.calculate_strain != 0;
    d.orderParameterCalculationIsOn = o.order_parameter != 0;
    d.A = o.A;
    d.KASQR = o.KASQR;
    d.fieldMode = o.field_mode;
    d.preconditioner = o.preconditioner;
    d.solverRelTolerance = o.solver_tol;
    d.solverMaxIterations = o.solver_max_iter;
    d.solverRestart = o.solver_restart;
    switch (o.linear_solver) {
    case SDDDST_LINSOLVE_UMFPACK: d.linearSolver = LinearSolver::Umfpack; break;
    case SDDDST_LINSOLVE_DENSE_LU: d.linearSolver = LinearSolver::DenseLU; break;
    case SDDDST_LINSOLVE_FGMRES: d.linearSolver = LinearSolver::FGMRES; break;
    default: d.linearSolver = LinearSolver::Auto; break;
    }
    d.tau.reset(new AnalyticField());
    if (o.protocol == SDDDST_PROTOCOL_FIXED_RATE) {
        FixedRateProtocol *p = new FixedRateProtocol();
        p->setRate(o.stress_param1);
        d.externalStressProtocol.reset(p);
    } else if (o.protocol == SDDDST_PROTOCOL_SPRING) {
        SpringProtocol *p = new SpringProtocol();
        p->setRate(o.stress_param1);
        p->setSpringConstant(o.stress_param2);
        d.externalStressProtocol.reset(p);
    } else if (o.protocol == SDDDST_PROTOCOL_CONSTANT) {
        d.externalStressProtocol.reset(new ConstantStressProtocol(o.stress_param1));
    } else {
        d.externalStressProtocol.reset(new StressProtocol());
    }
    if (o.log_path && *o.log_path) d.standardOutputLog = std::ofstream(o.log_path);
    if (o.spai_weights && *o.spai_weights) d.spaiWeightsPath = o.spai_weights;
}
}  // namespace

extern "C" {

void sdddst_sim_default_options(sdddst_sim_options *o)
{
    // Defaults of the reference's command line (project_parser.cpp:57-68).
    o->cutoff_multiplier = 1e20;
    o->precision = 1e-5;
    o->initial_step = 1e-6;
    o->max_step = -1;
    o->protocol = SDDDST_PROTOCOL_ZERO;
    o->stress_param1 = 0;
    o->stress_param2 = 1.0;
    o->calculate_strain = 0;
    o->order_parameter = 0;
    o->linear_solver = SDDDST_LINSOLVE_AUTO;
    o->solver_tol = 1e-10;
    o->solver_max_iter = 300;
    o->solver_restart = 40;
    o->field_mode = SDDDST_FIELD_FAST;
    o->preconditioner = SDDDST_PRECOND_PAIR;
    o->A = 1e-4 * 16.0;
    o->KASQR = 1.65 * 1.65 * 1e6 / 256.0;
    o->log_path = nullptr;
    o->spai_weights = nullptr;
}

int sdddst_sim_create(int n, const double *x, const double *y, const double *b, int pc, const double *px,
                      const double *py, const sdddst_sim_options *o, sdddst_sim **out)
{
    try {
        if (!o || !out || n < 2 || !x || !y || !b) {
            core_error = "bad arguments";
            return SDDDST_ERR_ARG;
        }
        std::unique_ptr<sdddst_sim> s(new sdddst_sim());
        s->sD = std::make_shared<SimulationData>();
        std::vector<Dislocation> d(n);
        for (int i = 0; i < n; ++i) d[i] = Dislocation{x[i], y[i], b[i]};
        s->sD->setDislocations(d);
        std::vector<PointDefect> p(pc > 0 ? pc : 0);
        for (int i = 0; i < pc; ++i) p[i] = PointDefect{px[i], py[i]};
        s->sD->setPointDefects(p);
        configure(*s->sD, *o);
        s->sim.reset(new Simulation(s->sD));
        *out = s.release();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_create_from_files(const char *dconf, const char *pdconf, const sdddst_sim_options *o,
                                 sdddst_sim **out)
{
    try {
        std::unique_ptr<sdddst_sim> s(new sdddst_sim());
        s->sD = std::make_shared<SimulationData>(dconf ? dconf : "", pdconf ? pdconf : "");
        configure(*s->sD, *o);
        s->sim.reset(new Simulation(s->sD));
        *out = s.release();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_destroy(sdddst_sim *s)
{
    if (s && s->owner && s->owner->sim == s) s->owner->sim = nullptr;
    delete s;
    return SDDDST_OK;
}

int sdddst_sim_step(sdddst_sim *s, int *success)
{
    try {
        const double t0 = get_wall_time();
        s->sim->applyDataChanges();
        s->lastSuccess = s->sim->step();
        s->lastWall = get_wall_time() - t0;
        if (success) *success = s->lastSuccess ? 1 : 0;
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_run(sdddst_sim *s, int successful)
{
    try {
        s->sD->isStepCountLimit = true;
        s->sD->stepCountLimit = (unsigned)(s->sD->succesfulSteps + successful);
        s->sim->applyDataChanges();
        s->sim->run();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_run_configured(sdddst_sim *s)
{
    try {
        s->sim->applyDataChanges();
        s->sim->run();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_set_step_size(sdddst_sim *s, double h)
{
    s->sD->stepSize = h;
    return SDDDST_OK;
}

int sdddst_sim_get_state(sdddst_sim *s, sdddst_sim_state *o)
{
    const SimulationData &d = *s->sD;
    o->sim_time = d.simTime;
    o->step_size = d.stepSize;
    o->strain = d.totalAccumulatedStrainIncrease;
    o->energy = s->sim->getEnergy();
    o->max_error_ratio_sqr = s->sim->precisionHandler().getMaxErrorRatioSqr();
    o->sum_avg_speed = d.sumAvgSpeed;
    o->cutoff = d.cutOff;
    o->succesful_steps = (long long)d.succesfulSteps;
    o->failed_steps = (long long)d.failedSteps;
    o->last_step_succeeded = s->lastSuccess ? 1 : 0;
    o->last_step_wall_s = s->lastWall;
    int it = 0, fb = 0;
    double rr = 0;
    int64_t ti = 0, ts = 0;
    sdddst_solver_stats(s->sim->deviceContext(), &it, &rr, &fb, &ti, &ts);
    o->solver_total_iters = ti;
    o->solver_total_solves = ts;
    o->solver_last_iters = it;
    o->solver_fallbacks = fb;
    o->solver_last_rel_residual = rr;
    int pf = 0;
    sdddst_solver_precond_fallbacks(s->sim->deviceContext(), &pf);
    o->solver_precond_fallbacks = pf;
    o->aborted_attempts = s->sim->abortedAttemptCount();
    int sr = 0, cnt = 0;
    // counts only (max_records 0): the records stay for sdddst_solver_history
    (void)sdddst_solver_history(s->sim->deviceContext(), nullptr, 0, &cnt, &sr);
    o->solver_strong_retries = sr;
    return SDDDST_OK;
}

int sdddst_sim_positions(sdddst_sim *s, double *x, int n)
{
    try {
        const std::vector<Dislocation> &d = s->sim->getStoredDislocationData();
        if (n != (int)d.size()) {
            core_error = "size mismatch";
            return SDDDST_ERR_ARG;
        }
        for (int i = 0; i < n; ++i) x[i] = d[i].x;
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_sim_flush_log(sdddst_sim *s)
{
    s->sD->standardOutputLog.flush();
    return SDDDST_OK;
}

int sdddst_sim_write_dconf(sdddst_sim *s, const char *path)
{
    try {
        s->sim->getStoredDislocationData();
        s->sD->writeDislocationDataToFile(path);
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

const char *sdddst_sim_solver_name(sdddst_sim *s)
{
    switch (s->sim->activeSolver()) {
    case LinearSolver::Umfpack: return "umfpack";
    case LinearSolver::DenseLU: return "dense-lu";
    case LinearSolver::FGMRES: return "fgmres";
    default: return "auto";
    }
}

const char *sdddst_core_last_error(void) { return core_error.c_str(); }

int sdddst_core_field_eval(int which, int n, const double *dx, const double *dy, double *out)
{
    try {
        if (which < 0 || which > 2 || n < 0 || (n > 0 && (!dx || !dy || !out))) {
            core_error = "bad arguments";
            return SDDDST_ERR_ARG;
        }
        AnalyticField f;
        Field &tau = f;
        for (int i = 0; i < n; ++i)
            out[i] = which == 0 ? tau.xy(dx[i], dy[i]) : which == 1 ? tau.xy_diff_x(dx[i], dy[i])
                                                                    : tau.xy_diff_x2(dx[i], dy[i]);
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

sdddst_ctx *sdddst_sim_device_context(sdddst_sim *s) { return s->sim->deviceContext(); }

// ---- SimulationData handles ----------------------------------------------

namespace {

// The SimulationData fields the bindings expose as properties
// (bindings.cpp:145-190), by the bindings' names.
struct Prop {
    const char *name;
    double SimulationData::*d;
    unsigned int SimulationData::*u;
    bool SimulationData::*b;
    size_t SimulationData::*z;
    bool readOnly;
};

const Prop kProps[] = {
    {"cutoff_multiplier", &SimulationData::cutOffMultiplier, nullptr, nullptr, nullptr, false},
    {"cutoff", &SimulationData::cutOff, nullptr, nullptr, nullptr, false},
    {"cutoff_square", &SimulationData::cutOffSqr, nullptr, nullptr, nullptr, false},
    {"one_per_cutoff_square", &SimulationData::onePerCutOffSqr, nullptr, nullptr, nullptr, false},
    {"precision", &SimulationData::prec, nullptr, nullptr, nullptr, false},
    {"point_defect_count", nullptr, &SimulationData::pc, nullptr, nullptr, true},
    {"dislocation_count", nullptr, &SimulationData::dc, nullptr, nullptr, true},
    {"iteration_count", nullptr, &SimulationData::ic, nullptr, nullptr, false},
    {"time_limit", &SimulationData::timeLimit, nullptr, nullptr, nullptr, false},
    {"step_size", &SimulationData::stepSize, nullptr, nullptr, nullptr, false},
    {"simulation_time", &SimulationData::simTime, nullptr, nullptr, nullptr, false},
    {"KASQR", &SimulationData::KASQR, nullptr, nullptr, nullptr, false},
    {"A", &SimulationData::A, nullptr, nullptr, nullptr, false},
    {"successful_steps", nullptr, nullptr, nullptr, &SimulationData::succesfulSteps, false},
    {"failed_steps", nullptr, nullptr, nullptr, &SimulationData::failedSteps, false},
    {"total_accumulated_strain_increase", &SimulationData::totalAccumulatedStrainIncrease, nullptr, nullptr, nullptr,
     false},
    {"strain_increase_limted", nullptr, nullptr, &SimulationData::isStrainIncreaseLimit, nullptr, false},
    {"total_accumulated_strain_increase_limit", &SimulationData::totalAccumulatedStrainIncreaseLimit, nullptr,
     nullptr, nullptr, false},
    {"max_step_size_limited", nullptr, nullptr, &SimulationData::isMaxStepSizeLimit, nullptr, false},
    {"max_stepsize_limit", &SimulationData::maxStepSizeLimit, nullptr, nullptr, nullptr, false},
    {"time_limited", nullptr, nullptr, &SimulationData::isTimeLimit, nullptr, false},
    {"step_count_limited", nullptr, nullptr, &SimulationData::isStepCountLimit, nullptr, false},
    {"step_count_limit", nullptr, &SimulationData::stepCountLimit, nullptr, nullptr, false},
    {"calculate_strain_during_simulation", nullptr, nullptr, &SimulationData::calculateStrainDuringSimulation,
     nullptr, false},
    {"calculate_order_parameter", nullptr, nullptr, &SimulationData::orderParameterCalculationIsOn, nullptr, false},
    {"count_avalanches", nullptr, nullptr, &SimulationData::countAvalanches, nullptr, false},
    {"avalanche_speed_threshold", &SimulationData::avalancheSpeedThreshold, nullptr, nullptr, nullptr, false},
    {"trigger_limit_for_avalanche_count", nullptr, &SimulationData::avalancheTriggerLimit, nullptr, nullptr, false},
    {"avalanche_count", nullptr, &SimulationData::avalancheCount, nullptr, nullptr, false},
    {"in_avalanche", nullptr, nullptr, &SimulationData::inAvalanche, nullptr, false},
    {"save_sub_configs", nullptr, nullptr, &SimulationData::isSaveSubConfigs, nullptr, false},
    {"sub_config_delay", nullptr, &SimulationData::subConfigDelay, nullptr, nullptr, false},
    {"sub_config_delay_during_avalanche", nullptr, &SimulationData::subConfigDelayDuringAvalanche, nullptr, nullptr,
     false},
    {"sub_config_distance_counter", nullptr, &SimulationData::subconfigDistanceCounter, nullptr, nullptr, false},
    // not in the bindings' list, but settable there through the same object
    {"speed_limited", nullptr, nullptr, &SimulationData::isSpeedLimit, nullptr, false},
    {"speed_limit", &SimulationData::speedLimit, nullptr, nullptr, nullptr, false},
    {"stress_limited", nullptr, nullptr, &SimulationData::isStressLimit, nullptr, false},
    {"stress_limit", &SimulationData::stressLimit, nullptr, nullptr, nullptr, false},
    {"remaining_final_steps", nullptr, &SimulationData::remainingFinalSteps, nullptr, nullptr, false},
};

const Prop *findProp(const char *name)
{
    if (!name) return nullptr;
    for (const Prop &p : kProps)
        if (std::strcmp(p.name, name) == 0) return &p;
    return nullptr;
}

int unknown(const char *name)
{
    core_error = std::string("SimulationData has no property '") + (name ? name : "(null)") + "'";
    return SDDDST_ERR_ARG;
}

std::vector<Dislocation> *dislocationVector(SimulationData &d, const char *which)
{
    const std::string w = which ? which : "";
    if (w == "dislocations") return &d.dislocations;
    if (w == "big_step") return &d.bigStep;
    if (w == "first_small_step") return &d.firstSmall;
    if (w == "second_small_step") return &d.secondSmall;
    return nullptr;
}

std::vector<double> *doubleVector(SimulationData &d, const char *which)
{
    const std::string w = which ? which : "";
    if (w == "g_vec") return &d.g;
    if (w == "init_speed") return &d.initSpeed;
    if (w == "init_speed_2") return &d.initSpeed2;
    if (w == "d_vec") return &d.dVec;
    return nullptr;
}

// Positions live on the device while a Simulation runs: bring them to the
// shared SimulationData before it is read.
void syncFromSimulation(sdddst_data *d)
{
    if (d->sim) d->sim->sim->getStoredDislocationData();
}

}  // namespace

int sdddst_data_create(const char *dconf, const char *pdconf, sdddst_data **out)
{
    try {
        if (!out) return SDDDST_ERR_ARG;
        std::unique_ptr<sdddst_data> d(new sdddst_data());
        d->sD = std::make_shared<SimulationData>(dconf ? dconf : "", pdconf ? pdconf : "");
        d->sD->tau.reset(new AnalyticField());
        d->sD->externalStressProtocol.reset(new StressProtocol());
        *out = d.release();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_data_destroy(sdddst_data *d)
{
    if (d && d->sim) d->sim->owner = nullptr;
    delete d;
    return SDDDST_OK;
}

int sdddst_data_get(sdddst_data *d, const char *name, double *value)
{
    const Prop *p = findProp(name);
    if (!p) return unknown(name);
    if (!d || !value) return SDDDST_ERR_ARG;
    SimulationData &s = *d->sD;
    if (p->d) *value = s.*(p->d);
    else if (p->u) *value = (double)(s.*(p->u));
    else if (p->b) *value = (s.*(p->b)) ? 1.0 : 0.0;
    else *value = (double)(s.*(p->z));
    return SDDDST_OK;
}

int sdddst_data_set(sdddst_data *d, const char *name, double value)
{
    const Prop *p = findProp(name);
    if (!p) return unknown(name);
    if (!d) return SDDDST_ERR_ARG;
    if (p->readOnly) {
        core_error = std::string("SimulationData property '") + name + "' is read-only";
        return SDDDST_ERR_ARG;
    }
    SimulationData &s = *d->sD;
    if (p->d) s.*(p->d) = value;
    else if (p->u) s.*(p->u) = (unsigned int)value;
    else if (p->b) s.*(p->b) = value != 0.0;
    else s.*(p->z) = (size_t)value;
    return SDDDST_OK;
}

int sdddst_data_get_string(sdddst_data *d, const char *name, char *buf, int cap)
{
    if (!d || !name || !buf || cap <= 0) return SDDDST_ERR_ARG;
    const std::string n = name;
    std::string v;
    if (n == "final_dislocation_configuration_path") v = d->sD->endDislocationConfigurationPath;
    else if (n == "sub_config_path") v = d->sD->subConfigPath;
    else return unknown(name);
    std::snprintf(buf, (size_t)cap, "%s", v.c_str());
    return (int)v.size() < cap ? SDDDST_OK : SDDDST_ERR_ARG;
}

int sdddst_data_set_string(sdddst_data *d, const char *name, const char *value)
{
    if (!d || !name || !value) return SDDDST_ERR_ARG;
    const std::string n = name;
    if (n == "final_dislocation_configuration_path") d->sD->endDislocationConfigurationPath = value;
    else if (n == "sub_config_path") d->sD->subConfigPath = value;
    else return unknown(name);
    return SDDDST_OK;
}

int sdddst_data_get_dislocations(sdddst_data *d, const char *which, double *x, double *y, double *b, int cap)
{
    try {
        if (!d) return SDDDST_ERR_ARG;
        if (std::string(which ? which : "") == "dislocations") syncFromSimulation(d);
        const std::vector<Dislocation> *v = dislocationVector(*d->sD, which);
        if (!v) return unknown(which);
        const int n = (int)v->size();
        for (int i = 0; i < n && i < cap; ++i) {
            if (x) x[i] = (*v)[i].x;
            if (y) y[i] = (*v)[i].y;
            if (b) b[i] = (*v)[i].b;
        }
        return n;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_data_set_dislocations(sdddst_data *d, const double *x, const double *y, const double *b, int n)
{
    try {
        if (!d || n < 0 || (n > 0 && (!x || !y || !b))) return SDDDST_ERR_ARG;
        std::vector<Dislocation> v(n);
        for (int i = 0; i < n; ++i) v[i] = Dislocation{x[i], y[i], b[i]};
        d->sD->setDislocations(v);
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_data_get_point_defects(sdddst_data *d, double *x, double *y, int cap)
{
    if (!d) return SDDDST_ERR_ARG;
    const int n = (int)d->sD->points.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (x) x[i] = d->sD->points[i].x;
        if (y) y[i] = d->sD->points[i].y;
    }
    return n;
}

int sdddst_data_set_point_defects(sdddst_data *d, const double *x, const double *y, int n)
{
    if (!d || n < 0 || (n > 0 && (!x || !y))) return SDDDST_ERR_ARG;
    std::vector<PointDefect> p(n);
    for (int i = 0; i < n; ++i) p[i] = PointDefect{x[i], y[i]};
    d->sD->setPointDefects(p);
    return SDDDST_OK;
}

int sdddst_data_get_vector(sdddst_data *d, const char *which, double *out, int cap)
{
    if (!d) return SDDDST_ERR_ARG;
    const std::vector<double> *v = doubleVector(*d->sD, which);
    if (!v) return unknown(which);
    const int n = (int)v->size();
    for (int i = 0; i < n && i < cap; ++i) out[i] = (*v)[i];
    return n;
}

int sdddst_data_set_vector(sdddst_data *d, const char *which, const double *in, int n)
{
    if (!d || n < 0 || (n > 0 && !in)) return SDDDST_ERR_ARG;
    std::vector<double> *v = doubleVector(*d->sD, which);
    if (!v) return unknown(which);
    v->assign(in, in + n);
    return SDDDST_OK;
}

int sdddst_data_file(sdddst_data *d, int op, const char *path)
{
    try {
        if (!d || !path) return SDDDST_ERR_ARG;
        switch (op) {
        case 0: d->sD->readDislocationDataFromFile(path); break;
        case 1:
            syncFromSimulation(d);
            d->sD->writeDislocationDataToFile(path);
            break;
        case 2: d->sD->readPointDefectDataFromFile(path); break;
        case 3: d->sD->writePointDefectDataToFile(path); break;
        default: core_error = "unknown file operation"; return SDDDST_ERR_ARG;
        }
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_data_init_simulation_variables(sdddst_data *d)
{
    if (!d) return SDDDST_ERR_ARG;
    d->sD->initSimulationVariables();
    return SDDDST_OK;
}

int sdddst_data_update_cutoff(sdddst_data *d)
{
    if (!d) return SDDDST_ERR_ARG;
    d->sD->updateCutOff();
    return SDDDST_OK;
}

int sdddst_data_set_field(sdddst_data *d, int kind)
{
    if (!d || kind < 0 || kind > 1) return SDDDST_ERR_ARG;
    if (kind == 1) d->sD->tau.reset(new AnalyticField());
    else d->sD->tau.reset(new Field());
    return SDDDST_OK;
}

int sdddst_data_field_eval(sdddst_data *d, int which, int n, const double *dx, const double *dy, double *out)
{
    try {
        if (!d || !d->sD->tau || which < 0 || which > 2 || n < 0 || (n > 0 && (!dx || !dy || !out)))
            return SDDDST_ERR_ARG;
        Field &tau = *d->sD->tau;
        for (int i = 0; i < n; ++i)
            out[i] = which == 0 ? tau.xy(dx[i], dy[i]) : which == 1 ? tau.xy_diff_x(dx[i], dy[i])
                                                                    : tau.xy_diff_x2(dx[i], dy[i]);
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int sdddst_data_set_protocol(sdddst_data *d, int kind, double p1, double p2)
{
    if (!d) return SDDDST_ERR_ARG;
    switch (kind) {
    case SDDDST_PROTOCOL_ZERO: d->sD->externalStressProtocol.reset(new StressProtocol()); break;
    case SDDDST_PROTOCOL_FIXED_RATE: {
        FixedRateProtocol *p = new FixedRateProtocol();
        p->setRate(p1);
        d->sD->externalStressProtocol.reset(p);
        break;
    }
    case SDDDST_PROTOCOL_SPRING: {
        SpringProtocol *p = new SpringProtocol();
        p->setRate(p1);
        p->setSpringConstant(p2);
        d->sD->externalStressProtocol.reset(p);
        break;
    }
    case SDDDST_PROTOCOL_CONSTANT: d->sD->externalStressProtocol.reset(new ConstantStressProtocol(p1)); break;
    default: core_error = "unknown stress protocol"; return SDDDST_ERR_ARG;
    }
    return SDDDST_OK;
}

int sdddst_data_protocol(sdddst_data *d, int op, double t, int type, double *out)
{
    try {
        if (!d || !d->sD->externalStressProtocol || type < 0 || type > 3) return SDDDST_ERR_ARG;
        StressProtocol &p = *d->sD->externalStressProtocol;
        const StressProtocolStepType st = (StressProtocolStepType)type;
        if (op == 1) {
            syncFromSimulation(d);
            p.calculateStress(t, d->sD->dislocations, st);
        }
        if (out) *out = p.getStress(st);
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

struct sdddst_ph {
    PrecisionHandler ph;
};

int sdddst_ph_create(sdddst_ph **out)
{
    if (!out) return SDDDST_ERR_ARG;
    *out = new sdddst_ph();
    return SDDDST_OK;
}

int sdddst_ph_destroy(sdddst_ph *p)
{
    delete p;
    return SDDDST_OK;
}

int sdddst_ph_call(sdddst_ph *p, int op, double a, unsigned int id, double *out)
{
    try {
        if (!p) return SDDDST_ERR_ARG;
        PrecisionHandler &h = p->ph;
        double r = 0;
        switch (op) {
        case 0: h.setSize((unsigned int)a); break;
        case 1: r = (double)h.getSize(); break;
        case 2: h.reset(); break;
        case 3: h.updateTolerance(a, id); break;
        case 4: h.updateError(a, id); break;
        case 5: r = h.getNewStepSize(a); break;
        case 6: r = h.getMinPrecisity(); break;
        case 7: h.setMinPrecisity(a); break;
        case 8: r = h.getMaxErrorRatioSqr(); break;
        default: core_error = "unknown PrecisionHandler operation"; return SDDDST_ERR_ARG;
        }
        if (out) *out = r;
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

double sdddst_core_utility(int which, double a, double b, double k)
{
    switch (which) {
    case 0: normalize(a); return a;
    case 1: return X(a);
    case 2: return X2(a);
    case 3: return E(a, b, k);
    case 4: return X_dx(a);
    case 5: return X2_dx(a);
    case 6: return E_dx(a, b, k);
    case 7: return get_wall_time();
    case 8: return get_cpu_time();
    default: return NAN;
    }
}

int sdddst_sim_create_from_data(sdddst_data *d, sdddst_sim **out)
{
    try {
        if (!d || !out) return SDDDST_ERR_ARG;
        std::unique_ptr<sdddst_sim> s(new sdddst_sim());
        s->sD = d->sD;
        s->sim.reset(new Simulation(s->sD));
        s->owner = d;
        d->sim = s.get();
        *out = s.release();
        return SDDDST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

}  // extern "C"
