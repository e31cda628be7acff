// Simulation (reference: src/simulation.cpp:38-668), driving the device
// through the C ABI.  The stage structure, the stress-protocol calls, the
// accept/reject bookkeeping and the 12-column log follow the reference line
// by line; the O(N^2) and O(nnz) work and the per-dislocation controller run
// on the MI355X, and the state stays resident there between calls.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "../../include/sdddst/StressProtocols/constant_stress_protocol.h"
#include "../../include/sdddst/StressProtocols/spring_protocol.h"
#include "../../include/sdddst/core_capi.h"
#include "../../include/sdddst/simulation.h"
#include "../../include/sdddst/utility.h"
#include "../../include/sdddst_amd.h"
#include "umfpack_loader.h"

using namespace sdddstCore;

namespace {
constexpr int DISL = SDDDST_SLOT_DISLOCATIONS;
constexpr int BIG = SDDDST_SLOT_BIG_STEP;
constexpr int FS = SDDDST_SLOT_FIRST_SMALL;
constexpr int SS = SDDDST_SLOT_SECOND_SMALL;
constexpr int SCRA = SDDDST_SLOT_SCRATCH_A;
constexpr int SCRB = SDDDST_SLOT_SCRATCH_B;
}  // namespace

void Simulation::check(int status, const char *what)
{
    if (status != SDDDST_OK) throw std::runtime_error(std::string(what) + ": " + sdddst_last_error());
}

Simulation::Simulation(std::shared_ptr<SimulationData> _sD)
    : succesfulStep(true), lastWriteTimeFinished(0), initSpeedCalculationIsNeeded(true), firstStepRequest(true),
      energy(0), energyAccum(0), vsquare(0), vsquare1(0), vsquare2(0), lastRunWall(0), sD(_sD),
      pH(new PrecisionHandler), ctx(nullptr), solver(LinearSolver::DenseLU), hostNnz(0)
{
    if (!sD) throw std::invalid_argument("Simulation: no SimulationData");
    if (!sD->tau || sD->tau->deviceKind() != 1)
        throw std::invalid_argument("Simulation: the device path implements the analytic periodic field only");
    if (!sD->externalStressProtocol) sD->externalStressProtocol.reset(new StressProtocol());
    sD->standardOutputLog << std::scientific << std::setprecision(16);
    pH->setMinPrecisity(sD->prec);
    pH->setSize(sD->dc);

    sdddst_params p{sD->cutOffMultiplier, sD->prec, sD->A, sD->KASQR, (int)sD->ic, sD->fieldMode};
    check(sdddst_ctx_create((int)sD->dc, (int)sD->pc, &p, &ctx), "sdddst_ctx_create");
    pushed = p;
    std::vector<double> x(sD->dc), y(sD->dc), b(sD->dc);
    for (unsigned i = 0; i < sD->dc; ++i) {
        x[i] = sD->dislocations[i].x;
        y[i] = sD->dislocations[i].y;
        b[i] = sD->dislocations[i].b;
    }
    check(sdddst_upload_state(ctx, DISL, x.data(), y.data(), b.data(), (int)sD->dc), "upload");
    if (sD->pc) {
        std::vector<double> px(sD->pc), py(sD->pc);
        for (unsigned i = 0; i < sD->pc; ++i) {
            px[i] = sD->points[i].x;
            py[i] = sD->points[i].y;
        }
        check(sdddst_upload_points(ctx, px.data(), py.data(), (int)sD->pc), "upload points");
    }
    // Solver resolution: UMFPACK when requested or available, device otherwise.
    LinearSolver want = sD->linearSolver;
    if (want == LinearSolver::Auto || want == LinearSolver::Umfpack) {
        std::string why;
        umf.reset(UmfpackLibrary::load(&why));
        if (umf) want = LinearSolver::Umfpack;
        else if (want == LinearSolver::Umfpack) throw std::runtime_error("UMFPACK requested: " + why);
        else want = LinearSolver::DenseLU;
    }
    solver = want;
    int kind = SDDDST_SOLVER_HOST;
    if (solver == LinearSolver::DenseLU) kind = SDDDST_SOLVER_DENSE_LU;
    if (solver == LinearSolver::FGMRES) kind = SDDDST_SOLVER_FGMRES;
    check(sdddst_solver_select(ctx, kind, sD->solverRelTolerance, sD->solverMaxIterations, sD->solverRestart),
          "solver select");
    check(sdddst_set_preconditioner(ctx, sD->preconditioner), "preconditioner");
    if (sD->preconditioner == SDDDST_PRECOND_LEARNED) loadSpaiWeights();
}

// Raw little-endian float32 weights of the learned preconditioner; the
// default file ships in data/ beside this library.
void Simulation::loadSpaiWeights()
{
    std::string path = sD->spaiWeightsPath;
    if (path.empty()) {
        Dl_info info;
        if (!dladdr((void *)&sdddst_sim_default_options, &info) || !info.dli_fname)
            throw std::runtime_error("cannot locate libsdddstCore.so for the default preconditioner weights");
        std::string lib = info.dli_fname;
        const size_t slash = lib.find_last_of('/');
        path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/data/spai_gnn_h32.f32";
    }
    const int count = sdddst_spai_param_count();
    std::vector<float> w(count);
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open preconditioner weights " + path);
    f.read(reinterpret_cast<char *>(w.data()), sizeof(float) * count);
    if (f.gcount() != (std::streamsize)(sizeof(float) * count) || f.peek() != std::char_traits<char>::eof())
        throw std::runtime_error("preconditioner weights " + path + " must hold exactly " + std::to_string(count) +
                                 " float32 values");
    check(sdddst_spai_set_weights(ctx, w.data(), count), "preconditioner weights");
}

Simulation::~Simulation()
{
    if (umf && sD->Numeric) umf->free_numeric(&sD->Numeric);
    if (ctx) sdddst_ctx_destroy(ctx);
}

void Simulation::pushParams()
{
    sdddst_params p{sD->cutOffMultiplier, sD->prec, sD->A, sD->KASQR, (int)sD->ic, sD->fieldMode};
    check(sdddst_set_params(ctx, &p), "set params");
    pushed = p;
}

// The SimulationData is shared with the caller (the bindings expose its
// fields as properties): settings changed since the last push reach the
// device context's copy and the precision handler before the next step.
void Simulation::applyDataChanges()
{
    const sdddst_params p{sD->cutOffMultiplier, sD->prec, sD->A, sD->KASQR, (int)sD->ic, sD->fieldMode};
    if (p.cutoff_multiplier == pushed.cutoff_multiplier && p.precision == pushed.precision && p.A == pushed.A &&
        p.KASQR == pushed.KASQR && p.iteration_count == pushed.iteration_count && p.field_mode == pushed.field_mode)
        return;
    if (p.precision != pushed.precision) pH->setMinPrecisity(sD->prec);
    sD->updateCutOff();
    pushParams();
}

void Simulation::uploadPositions(int slot, const std::vector<Dislocation> &in)
{
    std::vector<double> x(in.size());
    for (size_t i = 0; i < in.size(); ++i) x[i] = in[i].x;
    check(sdddst_upload_state(ctx, slot, x.data(), nullptr, nullptr, (int)in.size()), "upload");
}

void Simulation::syncPositionsToHost(int slot, std::vector<Dislocation> &out)
{
    std::vector<double> x(sD->dc);
    check(sdddst_download_x(ctx, slot, x.data(), (int)sD->dc), "download");
    out.resize(sD->dc);
    for (unsigned i = 0; i < sD->dc; ++i) {
        out[i].x = x[i];
        out[i].y = sD->dislocations[i].y;
        out[i].b = sD->dislocations[i].b;
    }
}

// The "zero-stress" shortcut of simulation.cpp:75; a constant stress takes it
// too (identical bits, see ConstantStressProtocol).
bool Simulation::shortcutFirstIteration() const
{
    const std::string t = sD->externalStressProtocol->getType();
    return t == "zero-stress" || t == "constant-stress";
}

void Simulation::setStress(double t, StressProtocolStepType type, double strain)
{
    if (sD->externalStressProtocol->getType() == "spring-stress")
        static_cast<SpringProtocol *>(sD->externalStressProtocol.get())->calculateStress(t, sD->dislocations, type,
                                                                                         strain);
    else
        sD->externalStressProtocol->calculateStress(t, sD->dislocations, type);
}

void Simulation::speedsSlot(int slot, int buf, bool updateTolerance)
{
    const double sigma = sD->externalStressProtocol->getStress(sD->currentStressStateType);
    check(sdddst_speeds(ctx, slot, buf, sigma, updateTolerance ? 1 : 0, nullptr), "speeds");
}

// ---------------------------------------------------------------------------
// simulation.cpp:133-193 on device slots.
void Simulation::calculateGSlots(double stepsize, int newSlot, int oldSlot, bool useSpeed2, bool calculateInitSpeed,
                                 bool useInitSpeedForFirstStep, StressProtocolStepType origin,
                                 StressProtocolStepType end)
{
    int isp = useSpeed2 ? SDDDST_V_INIT2 : SDDDST_V_INIT;
    int csp = useSpeed2 ? SDDDST_V_SPEED2 : SDDDST_V_SPEED;
    const bool spring = sD->externalStressProtocol->getType() == "spring-stress";
    if (calculateInitSpeed) {
        double t = sD->simTime;
        double tasi = sD->totalAccumulatedStrainIncrease;
        if (origin == EndOfFirstSmallStep) {
            t += sD->stepSize * 0.5;
            if (spring) {
                double inc = 0;
                check(sdddst_strain_increment(ctx, DISL, FS, &inc), "strain");
                tasi += inc;
            }
        }
        setStress(t, origin, tasi);
        sD->currentStressStateType = origin;
        speedsSlot(oldSlot, isp, true);
    }
    if (useInitSpeedForFirstStep) {
        csp = isp;
    } else {
        double t = sD->simTime + sD->stepSize;
        double tasi = sD->totalAccumulatedStrainIncrease;
        if (spring) {
            double a = 0, b = 0;
            if (end == EndOfFirstSmallStep) {
                check(sdddst_strain_increment(ctx, DISL, FS, &a), "strain");
                tasi += a;
            } else if (end == EndOfSecondSmallStep) {
                check(sdddst_strain_increment(ctx, DISL, FS, &a), "strain");
                check(sdddst_strain_increment(ctx, FS, SS, &b), "strain");
                tasi += a + b;
            } else if (end == EndOfBigStep) {
                check(sdddst_strain_increment(ctx, DISL, BIG, &a), "strain");
                tasi += a;
            }
        }
        if (end == EndOfFirstSmallStep) t -= sD->stepSize * 0.5;
        setStress(t, end, tasi);
        sD->currentStressStateType = end;
        speedsSlot(newSlot, csp, true);
    }
    // g stays on the device for the device solvers; UMFPACK reads it on the host
    check(sdddst_residual(ctx, stepsize, newSlot, oldSlot, csp, isp,
                          solver == LinearSolver::Umfpack ? sD->g.data() : nullptr),
          "residual");
}

// simulation.cpp:62-84 on device slots.
// samePositions: the Jacobian is assembled at the positions of the previous
// assembly (stage II starts where stage I started), so the sparsity pattern
// and the field derivatives are reused and only the h-dependent values change.
void Simulation::integrateSlots(double stepsize, int newSlot, int oldSlot, bool useSpeed2, bool calculateInitSpeed,
                                StressProtocolStepType origin, StressProtocolStepType end, bool samePositions)
{
    check(sdddst_jacobian_csc(ctx, newSlot, stepsize, samePositions ? 2 : 1, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr),
          "jacobian");
    calculateSparseFormForJacobian();
    for (size_t i = 0; i < sD->ic; i++) {
        if (i > 0)
            calculateGSlots(stepsize, newSlot, oldSlot, useSpeed2, false, false, origin, end);
        else
            calculateGSlots(stepsize, newSlot, oldSlot, useSpeed2, calculateInitSpeed, shortcutFirstIteration(),
                            origin, end);
        solveEQSys();
        check(sdddst_newton_update(ctx, newSlot, solver == LinearSolver::Umfpack ? sD->x : nullptr), "update");
    }
    if (solver == LinearSolver::Umfpack && sD->Numeric) umf->free_numeric(&sD->Numeric);
}

// The factorisation of simulation.cpp:386-391: host UMFPACK on the CSC
// downloaded into SimulationData's arrays, or the device solver.
void Simulation::calculateSparseFormForJacobian()
{
    if (solver != LinearSolver::Umfpack) {
        check(sdddst_solver_factor(ctx), "factor");
        return;
    }
    int64_t nnz = 0;
    check(sdddst_jacobian_csc(ctx, DISL, 0.0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, &nnz), "nnz");
    if (nnz > hostNnz) {
        free(sD->Ai);
        free(sD->Ax);
        sD->Ai = (int *)malloc(sizeof(int) * nnz);
        sD->Ax = (double *)malloc(sizeof(double) * nnz);
        if (!sD->Ai || !sD->Ax) throw std::bad_alloc();
        hostNnz = nnz;
        sD->currentStorageSize = (unsigned)nnz;
    }
    check(sdddst_jacobian_csc(ctx, DISL, 0.0, 0, sD->Ap, sD->Ai, sD->Ax, sD->dVec.data(), sD->indexes.data(), &nnz),
          "download CSC");
    (void)umf->symbolic((int)sD->dc, (int)sD->dc, sD->Ap, sD->Ai, sD->Ax, &sD->Symbolic);
    (void)umf->numeric(sD->Ap, sD->Ai, sD->Ax, sD->Symbolic, &sD->Numeric);
    umf->free_symbolic(&sD->Symbolic);
}

// simulation.cpp:393-396
namespace {
// A Newton solve the device solver could not bring to its tolerance with no
// direct solver allowed at this size: the attempt ends there.
struct AttemptAborted {
};
}  // namespace

void Simulation::solveEQSys()
{
    if (solver != LinearSolver::Umfpack) {
        const int rc = sdddst_solver_solve(ctx, nullptr);
        if (rc == SDDDST_ERR_NOT_CONVERGED) throw AttemptAborted();
        check(rc, "solve");
        return;
    }
    (void)umf->solve(sD->Ap, sD->Ai, sD->Ax, sD->x, sD->g.data(), sD->Numeric);
}

void Simulation::calculateXError()
{
    // Folded into the device error controller (see stepStageIII).
}

double Simulation::calculateOrderParameter(const std::vector<double> &speeds)
{
    double orderParameter = 0;
    for (size_t i = 0; i < sD->dc; i++) orderParameter += sD->dislocations[i].b * speeds[i];
    return orderParameter;
}

// simulation.cpp:195-229 over the host CSC arrays.
double Simulation::getElement(int j, int si, int ei)
{
    int len = ei - si;
    if (len > 1) {
        int tmp = len / 2;
        double a;
        if (sD->Ai[si + tmp] > j) {
            a = getElement(j, si, si + tmp);
            if (a != 0.0) return a;
        } else {
            a = getElement(j, si + tmp, ei);
            if (a != 0.0) return a;
        }
    } else {
        if (sD->Ai[si] == j) return sD->Ax[si];
        return 0;
    }
    return 0;
}

double Simulation::getSimTime() { return sD->simTime; }

// ---------------------------------------------------------------------------
// Host-vector forms of the reference's public methods.

void Simulation::calculateSpeeds(const std::vector<Dislocation> &dis, std::vector<double> &res, bool ignorePHUpdate)
{
    uploadPositions(SCRA, dis);
    speedsSlot(SCRA, SDDDST_V_SCRATCH, !ignorePHUpdate);
    res.resize(sD->dc);
    check(sdddst_download_speed(ctx, SDDDST_V_SCRATCH, res.data(), (int)sD->dc), "download speeds");
}

void Simulation::calculateJacobian(const double &stepsize, const std::vector<Dislocation> &data)
{
    uploadPositions(SCRA, data);
    int64_t nnz = 0;
    check(sdddst_jacobian_csc(ctx, SCRA, stepsize, 1, nullptr, nullptr, nullptr, nullptr, nullptr, &nnz),
          "jacobian");
    if (nnz > hostNnz) {
        free(sD->Ai);
        free(sD->Ax);
        sD->Ai = (int *)malloc(sizeof(int) * nnz);
        sD->Ax = (double *)malloc(sizeof(double) * nnz);
        if (!sD->Ai || !sD->Ax) throw std::bad_alloc();
        hostNnz = nnz;
        sD->currentStorageSize = (unsigned)nnz;
    }
    check(sdddst_jacobian_csc(ctx, SCRA, stepsize, 0, sD->Ap, sD->Ai, sD->Ax, sD->dVec.data(), sD->indexes.data(),
                              &nnz),
          "download CSC");
}

void Simulation::calculateG(const double &stepsize, std::vector<Dislocation> &newDislocation,
                            const std::vector<Dislocation> &old, bool useSpeed2, bool calculateInitSpeed,
                            bool useInitSpeedForFirstStep, StressProtocolStepType origin, StressProtocolStepType end)
{
    uploadPositions(SCRA, newDislocation);
    uploadPositions(SCRB, old);
    calculateGSlots(stepsize, SCRA, SCRB, useSpeed2, calculateInitSpeed, useInitSpeedForFirstStep, origin, end);
    const int isp = useSpeed2 ? SDDDST_V_INIT2 : SDDDST_V_INIT;
    const int csp = useInitSpeedForFirstStep ? isp : (useSpeed2 ? SDDDST_V_SPEED2 : SDDDST_V_SPEED);
    check(sdddst_residual(ctx, stepsize, SCRA, SCRB, csp, isp, sD->g.data()), "residual");
}

void Simulation::integrate(const double &stepsize, std::vector<Dislocation> &newDislocation,
                           const std::vector<Dislocation> &old, bool useSpeed2, bool calculateInitSpeed,
                           StressProtocolStepType origin, StressProtocolStepType end)
{
    uploadPositions(SCRA, newDislocation);
    uploadPositions(SCRB, old);
    integrateSlots(stepsize, SCRA, SCRB, useSpeed2, calculateInitSpeed, origin, end);
    std::vector<double> x(sD->dc);
    check(sdddst_download_x(ctx, SCRA, x.data(), (int)sD->dc), "download");
    for (unsigned i = 0; i < sD->dc; ++i) newDislocation[i].x = x[i];
}

const std::vector<Dislocation> &Simulation::getStoredDislocationData()
{
    syncPositionsToHost(DISL, sD->dislocations);
    return sD->dislocations;
}

// ---------------------------------------------------------------------------
// simulation.cpp:417-433
void Simulation::run()
{
    const double t0 = get_wall_time();
    const size_t s0 = sD->succesfulSteps;
    while (((sD->isTimeLimit && sD->simTime < sD->timeLimit) || !sD->isTimeLimit) &&
           ((sD->isStrainIncreaseLimit &&
             sD->totalAccumulatedStrainIncrease < sD->totalAccumulatedStrainIncreaseLimit) ||
            !sD->isStrainIncreaseLimit) &&
           ((sD->isStepCountLimit && sD->succesfulSteps < sD->stepCountLimit) || !sD->isStepCountLimit) &&
           ((sD->countAvalanches && sD->avalancheCount < sD->avalancheTriggerLimit) || !sD->countAvalanches) &&
           !sD->finish) {
        step();
    }
    while (sD->remainingFinalSteps > 0) {
        sD->inFinal = true;
        step();
    }
    const size_t done = sD->succesfulSteps - s0;
    lastRunWall = done ? (get_wall_time() - t0) / (double)done : 0.0;
    syncPositionsToHost(DISL, sD->dislocations);
    if (!sD->endDislocationConfigurationPath.empty())
        sD->writeDislocationDataToFile(sD->endDislocationConfigurationPath);
}

// simulation.cpp:435-441
bool Simulation::step()
{
    try {
        stepStageI();
        stepStageII();
        stepStageIII();
    } catch (const AttemptAborted &) {
        rejectAbortedAttempt();
    }
    return succesfulStep;
}

// An attempt whose linear solve did not converge (the reference's UMFPACK
// would have solved it; here FGMRES stagnates on the indefinite system of a
// far too large step and N is above the dense solver's limit) is rejected
// like a step over the error bound (simulation.cpp:638-649), with the step
// size divided by 8 since no error estimate exists.  The positions of the
// attempt are discarded; the state of the last accepted step is untouched.
void Simulation::rejectAbortedAttempt()
{
    succesfulStep = false;
    sD->failedSteps++;
    abortedAttempts++;
    check(sdddst_reset_tolerance(ctx), "reset tolerance");
    pH->reset();
    sD->stepSize *= 0.125;
    if (sD->isMaxStepSizeLimit && sD->maxStepSizeLimit < sD->stepSize) sD->stepSize = sD->maxStepSizeLimit;
}

void Simulation::writeLogLine(const std::string &line) { sD->standardOutputLog << line; }

// simulation.cpp:443-492
void Simulation::stepStageI()
{
    energyAccum = 0;
    double sumAvgSp = 0;
    vsquare = 0;
    sD->currentStressStateType = Original;
    if (firstStepRequest) {
        lastWriteTimeFinished = get_wall_time();
        setStress(sD->simTime, Original, sD->totalAccumulatedStrainIncrease);
        speedsSlot(DISL, SDDDST_V_INIT, true);
        initSpeedCalculationIsNeeded = false;
        double sa = 0, sq = 0;
        check(sdddst_speed_sums(ctx, SDDDST_V_INIT, &sa, &sq, nullptr), "sums");
        sumAvgSp = sa / double(sD->dc);
        vsquare = sq;
        std::ostringstream l;
        l << std::scientific << std::setprecision(16);
        l << sD->simTime << " " << sD->succesfulSteps << " " << sD->failedSteps << " " << 0 << " " << sumAvgSp << " "
          << sD->cutOff << " "
          << "-"
          << " " << sD->externalStressProtocol->getStress(sD->currentStressStateType) << " "
          << "-"
          << " " << sD->totalAccumulatedStrainIncrease << " " << vsquare << " " << energy << "\n";
        writeLogLine(l.str());
        firstStepRequest = false;
    }
    check(sdddst_copy_slot(ctx, BIG, DISL), "copy");
    integrateSlots(sD->stepSize, BIG, DISL, false, initSpeedCalculationIsNeeded, Original, EndOfBigStep);
    succesfulStep = false;
}

// simulation.cpp:494-499
void Simulation::stepStageII()
{
    check(sdddst_copy_slot(ctx, FS, DISL), "copy");
    integrateSlots(0.5 * sD->stepSize, FS, DISL, false, false, Original, EndOfFirstSmallStep, true);
}

// simulation.cpp:501-650
void Simulation::stepStageIII()
{
    sD->sumAvgSpeed = 0;
    check(sdddst_copy_slot(ctx, SS, FS), "copy");
    integrateSlots(0.5 * sD->stepSize, SS, FS, true, true, EndOfFirstSmallStep, EndOfSecondSmallStep);

    // the energy sums, calculateXError + the controller (the device also
    // resets the tolerance) and the strain increments, in one round trip
    int accept = 0, sel = 0;
    double hnext = 0, maxRatio = 0, strainA = 0, strainB = 0;
    const double hmax = sD->isMaxStepSizeLimit ? sD->maxStepSizeLimit : -1.0;
    check(sdddst_stage_end(ctx, sD->stepSize, hmax, sD->calculateStrainDuringSimulation ? 1 : 0, &vsquare1,
                           &vsquare2, &accept, &hnext, &maxRatio, &sel, &strainA, &strainB),
          "stage end");
    energyAccum = (vsquare1 + vsquare2) * 0.5 * sD->stepSize * 0.5;
    pH->setFromDevice(maxRatio, (unsigned)sel);

    if (pH->getMaxErrorRatioSqr() < 1.0) {
        succesfulStep = true;
        initSpeedCalculationIsNeeded = true;
        if (sD->inFinal) sD->remainingFinalSteps--;
        if (sD->calculateStrainDuringSimulation) {
            sD->totalAccumulatedStrainIncrease += strainA;
            sD->totalAccumulatedStrainIncrease += strainB;
        }
        check(sdddst_swap_slots(ctx, DISL, SS), "swap");
        sD->simTime += sD->stepSize;
        sD->succesfulSteps++;

        double orderParameter = 0;
        if (sD->orderParameterCalculationIsOn)
            check(sdddst_speed_sums(ctx, SDDDST_V_SPEED, nullptr, nullptr, &orderParameter), "order parameter");

        double current_wall_time = get_wall_time();

        sD->currentStressStateType = Original;
        setStress(sD->simTime, Original, sD->totalAccumulatedStrainIncrease);
        // the reference's tolerance updates here are discarded by the reset at
        // simulation.cpp:644, which the device controller already applied
        speedsSlot(DISL, SDDDST_V_INIT, false);
        initSpeedCalculationIsNeeded = false;
        double sa = 0, sq = 0;
        check(sdddst_speed_sums(ctx, SDDDST_V_INIT, &sa, &sq, nullptr), "sums");
        sD->sumAvgSpeed = sa / double(sD->dc);

        if (sD->sumAvgSpeed < sD->speedLimit && sD->isSpeedLimit) sD->finish = true;

        vsquare = sq;

        if (sD->countAvalanches) {
            if (sD->inAvalanche && sD->sumAvgSpeed < sD->avalancheSpeedThreshold) {
                sD->avalancheCount++;
                sD->inAvalanche = false;
            } else if (sD->sumAvgSpeed > sD->avalancheSpeedThreshold) {
                sD->inAvalanche = true;
            }
        }

        energyAccum += (vsquare2 + vsquare) * 0.5 * sD->stepSize * 0.5;
        std::ostringstream l;
        l << std::scientific << std::setprecision(16);
        l << sD->simTime << " " << sD->succesfulSteps << " " << sD->failedSteps << " " << pH->getMaxErrorRatioSqr()
          << " " << sD->sumAvgSpeed << " " << sD->cutOff << " ";
        if (sD->orderParameterCalculationIsOn) l << orderParameter;
        else l << "-";
        l << " " << sD->externalStressProtocol->getStress(Original) << " "
          << current_wall_time - lastWriteTimeFinished;
        if (sD->calculateStrainDuringSimulation) l << " " << sD->totalAccumulatedStrainIncrease;
        else l << " -";

        if (sD->isStressLimit && sD->externalStressProtocol->getStress(Original) > sD->stressLimit) sD->finish = true;

        if (sD->isSpeedThresholdForCutoffChange && sD->speedThresholdForCutoffChange > sD->sumAvgSpeed) {
            sD->cutOffMultiplier = 1e20;
            sD->updateCutOff();
            pushParams();
        }

        energy += energyAccum;
        l << " " << vsquare << " " << energy;
        l << "\n";
        writeLogLine(l.str());

        if (sD->isSaveSubConfigs) {
            if ((!sD->inAvalanche && sD->subConfigDelay >= sD->subconfigDistanceCounter) ||
                (sD->inAvalanche && sD->subConfigDelayDuringAvalanche >= sD->subconfigDistanceCounter)) {
                sD->subconfigDistanceCounter = 0;
                std::stringstream ss;
                ss << std::setprecision(16);
                ss << sD->simTime;
                syncPositionsToHost(DISL, sD->dislocations);
                sD->writeDislocationDataToFile(sD->subConfigPath + "/" + ss.str() + ".dconf");
            } else {
                sD->subconfigDistanceCounter++;
            }
        }

        lastWriteTimeFinished = get_wall_time();
    } else {
        sD->failedSteps++;
    }

    // getNewStepSize + clamp (device), reset (device + here)
    sD->stepSize = hnext;
    pH->reset();
}
