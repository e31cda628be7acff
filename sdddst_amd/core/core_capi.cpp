// C entry points over sdddstCore::Simulation (include/sdddst/core_capi.h).
#include <exception>
#include <memory>
#include <string>

#include "../../include/sdddst/Fields/AnalyticField.h"
#include "../../include/sdddst/StressProtocols/constant_stress_protocol.h"
#include "../../include/sdddst/StressProtocols/fixed_rate_protocol.h"
#include "../../include/sdddst/StressProtocols/spring_protocol.h"
#include "../../include/sdddst/core_capi.h"
#include "../../include/sdddst/simulation.h"
#include "../../include/sdddst/utility.h"
#include "../../include/sdddst_amd.h"

using namespace sdddstCore;

struct sdddst_sim {
    std::shared_ptr<SimulationData> sD;
    std::unique_ptr<Simulation> sim;
    bool lastSuccess = false;
    double lastWall = 0;
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
    delete s;
    return SDDDST_OK;
}

int sdddst_sim_step(sdddst_sim *s, int *success)
{
    try {
        const double t0 = get_wall_time();
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

}  // extern "C"
