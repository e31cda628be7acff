/*
 * sdddst_amd: C ABI of the MI355X (gfx950) implicit-step hot path of the
 * SDDDST 2-D periodic discrete dislocation dynamics simulator.
 *
 * The reference (pgabor/sdddst) has no FFI on this path: its boundary is the
 * C++ class API of sdddstCore::Simulation / Field / SimulationData.  The
 * entry points below are what those methods bind to in this build; each one
 * names the reference code it replaces.  The C++ mirror of that API
 * (include/sdddst/simulation.h and friends) calls only these functions.
 *
 * Conventions
 *   - Every function returns an int status: SDDDST_OK (0) or a negative
 *     SDDDST_ERR_* code; sdddst_last_error() gives the message of the last
 *     failure on the calling thread.  Nothing throws across the ABI.
 *   - FP64 everywhere on the path; indices are int32 (UMFPACK "di").
 *   - Host buffers are caller-owned; device buffers are context-owned.
 *   - One HIP stream per context; a context is not thread-safe.  Calls that
 *     return host data synchronise the context's stream.
 *   - Position slots follow SimulationData (simulation_data.h:69-127):
 *     0 dislocations, 1 bigStep, 2 firstSmall, 3 secondSmall, plus two
 *     scratch slots.  Only x moves (simulation.cpp:80); y and b are shared.
 *   - Speed buffers: 0 initSpeed, 1 initSpeed2, 2 speed, 3 speed2
 *     (simulation_data.h:75-81), 4 scratch.
 */
#ifndef SDDDST_AMD_H
#define SDDDST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDDDST_OK 0
#define SDDDST_ERR_ARG (-1)      /* bad argument or size                 */
#define SDDDST_ERR_HIP (-2)      /* HIP runtime / device failure          */
#define SDDDST_ERR_STATE (-3)    /* call out of order (e.g. no Jacobian)  */
#define SDDDST_ERR_SOLVER (-4)   /* factorisation / solve failure         */
#define SDDDST_ERR_NOMEM (-5)    /* device allocation failed              */
#define SDDDST_ERR_UNSUPPORTED (-6)
#define SDDDST_ERR_NOT_CONVERGED (-7) /* iterative solve above the direct solver's size limit
                                         did not reach its tolerance; delta holds the last
                                         iterate and the state is otherwise intact */

#define SDDDST_SLOT_DISLOCATIONS 0
#define SDDDST_SLOT_BIG_STEP 1
#define SDDDST_SLOT_FIRST_SMALL 2
#define SDDDST_SLOT_SECOND_SMALL 3
#define SDDDST_SLOT_SCRATCH_A 4    /* for host-vector API calls */
#define SDDDST_SLOT_SCRATCH_B 5
#define SDDDST_NUM_SLOTS 6

#define SDDDST_V_INIT 0
#define SDDDST_V_INIT2 1
#define SDDDST_V_SPEED 2
#define SDDDST_V_SPEED2 3
#define SDDDST_V_SCRATCH 4
#define SDDDST_NUM_VBUFS 5

/* Field evaluation modes of the pair-speed kernel. */
#define SDDDST_FIELD_FAST 0     /* near image as the reference, far images from one shared exp */
#define SDDDST_FIELD_EXACT 1    /* every image as the reference (bit-compatible per pair term) */
#define SDDDST_FIELD_ORDERED 2  /* exact terms summed in the reference's order (validation; 2x pair work) */

/* Linear solvers for the Newton correction. */
#define SDDDST_SOLVER_DENSE_LU 0   /* device LU with partial pivoting (rocSOLVER getrf/getrs) */
#define SDDDST_SOLVER_FGMRES 1     /* device FGMRES, FP64, relative residual tolerance        */
#define SDDDST_SOLVER_HOST 2       /* host hand-off: CSC to the caller (UMFPACK)              */

/* FGMRES right preconditioners. */
#define SDDDST_PRECOND_JACOBI 0   /* 1 / A_ii                                              */
#define SDDDST_PRECOND_PAIR 1     /* exact inverse of 2x2 blocks of mutually strongest pairs */
#define SDDDST_PRECOND_LEARNED 2  /* sparse approximate inverse on A's pattern from the net  */
#define SDDDST_PRECOND_STRONG 3   /* pair blocks + dense LU of the rows with |A_ii - 1| or an
                                     off-diagonal |A_ij| above 0.25 (at most 32768 rows);
                                     also the retry of any solve the others do not converge */

typedef struct sdddst_ctx sdddst_ctx;

/* Scalars of SimulationData that the kernels need
 * (simulation_data.h:84-107, constants.h:23-35). */
typedef struct sdddst_params {
    double cutoff_multiplier; /* cutOffMultiplier; rc = m/sqrt(N), m := 1e20 if m >= sqrt(2N)/12 */
    double precision;         /* prec: tolerance floor prec^2                                     */
    double A;                 /* point-defect interaction strength                                */
    double KASQR;             /* point-defect range parameter                                     */
    int iteration_count;      /* ic: fixed Newton iterations per integrate (2)                    */
    int field_mode;           /* SDDDST_FIELD_*                                                   */
} sdddst_params;

/* Library version string and last error message (thread-local). */
const char *sdddst_version(void);
const char *sdddst_last_error(void);

/* Context lifetime.  Replaces SimulationData's buffer allocation
 * (simulation_data.cpp:277-298); device memory is O(nnz), not O(N^2). */
int sdddst_ctx_create(int n, int pc, const sdddst_params *params, sdddst_ctx **out);
int sdddst_ctx_destroy(sdddst_ctx *ctx);
int sdddst_set_params(sdddst_ctx *ctx, const sdddst_params *params);

/* Cutoff of SimulationData::updateCutOff (simulation_data.cpp:208-218). */
int sdddst_get_cutoff(sdddst_ctx *ctx, double *cutoff, double *cutoff_sqr);

/* Positions.  upload: x (and y, b when non-NULL) of a slot; y and b are
 * shared.  download_x: x of a slot.  copy_slot: x[dst] = x[src]
 * (bigStep = dislocations etc., simulation.cpp:482, 496, 504). */
int sdddst_upload_state(sdddst_ctx *ctx, int slot, const double *x, const double *y, const double *b, int n);
int sdddst_download_x(sdddst_ctx *ctx, int slot, double *x, int n);
int sdddst_copy_slot(sdddst_ctx *ctx, int dst, int src);
int sdddst_swap_slots(sdddst_ctx *ctx, int a, int b);
int sdddst_upload_points(sdddst_ctx *ctx, const double *px, const double *py, int pc);

/* Simulation::calculateSpeeds (simulation.cpp:86-131) with
 * PrecisionHandler::updateTolerance (precision_handler.cpp:72-87) folded in:
 * v[out_buf] = pair speeds of slot cfg_slot + point-defect terms + b*sigma.
 * update_tol != 0 lowers the device tolerance as the reference does.
 * v_out (nullable) receives a host copy. */
int sdddst_speeds(sdddst_ctx *ctx, int cfg_slot, int out_buf, double sigma, int update_tol, double *v_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): v = the pair speeds of n dislocations plus b*sigma, as
 * sdddst_speeds computes them without point defects and without a
 * tolerance update, from caller-owned DEVICE arrays x, y, b (n each) into
 * the caller-owned device array v (n), enqueued on `stream` (a hipStream_t;
 * NULL: the null stream); returns once enqueued.  field_mode is
 * SDDDST_FIELD_FAST or SDDDST_FIELD_EXACT.  No context: partial sums and
 * per-dislocation factors live in a static device array of the library (a
 * fixed 8.6 MiB per GPU, held until the process ends), so n <= 8192
 * (SDDDST_ERR_ARG above).  The call allocates nothing; every call on one GPU
 * shares that array, so calls that overlap on one GPU race.  Tiles follow
 * index order, so results equal sdddst_speeds' to rounding, not bit for bit. */
int sdddst_speeds_noalloc(const double *x, const double *y, const double *b, int n, double sigma, int field_mode,
                          double *v, void *stream);

/* No counterpart in the reference: a measurement helper.  The FP64 VALU
 * rate this device sustains on independent FMA and multiply streams (8
 * chains of `iters` operations per lane, 16 waves per CU, the best of three
 * timed launches after a warm-up): wave-level instructions per second and
 * the FMA stream's TFLOP/s.  Beside the published FP64 peak, which assumes
 * the peak engine clock, this is the ceiling the pair kernel's VALU rate is
 * compared with.  Allocates a small output buffer and a stream of its own. */
int sdddst_fp64_probe(int iters, double *fma_tflops, double *fma_valu_per_s, double *mul_valu_per_s);

/* PrecisionHandler state: tolerance download / reset (precision_handler.cpp:62-70). */
int sdddst_get_tolerance(sdddst_ctx *ctx, double *tol_out, int n);
int sdddst_reset_tolerance(sdddst_ctx *ctx);
int sdddst_download_speed(sdddst_ctx *ctx, int buf, double *v_out, int n);
int sdddst_upload_speed(sdddst_ctx *ctx, int buf, const double *v, int n);
int sdddst_copy_speed(sdddst_ctx *ctx, int dst, int src);

/* Simulation::calculateJacobian (simulation.cpp:236-384): the CSC of
 * A = I + diag((1+d)/2) * M for slot cfg_slot and step h, built on the
 * device with the reference's exact sparsity pattern.  rebuild = 1 builds
 * it for (cfg_slot, h); rebuild = 2 keeps the sparsity pattern and the
 * per-pair field derivatives of the last rebuild = 1 (the caller asserts the
 * positions are the same, as for stages I and II of a step) and recomputes
 * the h-dependent values; rebuild = 0 reuses the last matrix.  Two-phase: with
 * Ap == NULL only *nnz is returned; a call with non-NULL pointers copies
 * Ap[n+1], Ai[nnz], Ax[nnz], and (each nullable) dVec[n] and the diagonal
 * positions diag[n] (SimulationData::indexes) to the host. */
int sdddst_jacobian_csc(sdddst_ctx *ctx, int cfg_slot, double h, int rebuild, int *Ap, int *Ai, double *Ax,
                        double *dVec, int *diag, int64_t *nnz);

/* Simulation::calculateG (simulation.cpp:189-192): g = x_new -
 * (1+d)/2*h*v[csp_buf] - x_old - (1-d)/2*h*v[isp_buf]; g_out nullable. */
int sdddst_residual(sdddst_ctx *ctx, double h, int new_slot, int old_slot, int csp_buf, int isp_buf, double *g_out);

/* Device linear solve of A delta = g with the current Jacobian
 * (replaces calculateSparseFormForJacobian + solveEQSys,
 * simulation.cpp:386-396, when the solver is a device solver).
 * factor: once per Jacobian; solve: delta on the device, delta_out nullable. */
int sdddst_solver_select(sdddst_ctx *ctx, int solver, double rel_tol, int max_iter, int restart);
int sdddst_set_preconditioner(sdddst_ctx *ctx, int precond);
int sdddst_solver_factor(sdddst_ctx *ctx);
int sdddst_solver_solve(sdddst_ctx *ctx, double *delta_out);
int sdddst_solver_stats(sdddst_ctx *ctx, int *last_iters, double *last_rel_residual, int *fallbacks,
                        int64_t *total_iters, int64_t *total_solves);
/* Learned-preconditioned solves that did not converge within one restart
 * cycle and were solved again with the pair-block preconditioner (no
 * counterpart in the reference). */
int sdddst_solver_precond_fallbacks(sdddst_ctx *ctx, int *count);
/* The most recent iterative-solver attempts (up to 1024 kept, oldest first,
 * at most max_records returned; a call with max_records > 0 clears the log,
 * one with 0 only reads the counts): seven
 * doubles per attempt {preconditioner (-1: the direct solver), iterations,
 * relative residual, ||g||, outcome (0 converged, 1 not converged, 2 direct
 * solve after it), h of the Jacobian, rows of the strong-row block (0 for
 * other preconditioners)}; strong (nullable) receives the count of
 * solves retried with the strong-row block preconditioner
 * (SDDDST_PRECOND_STRONG) before the direct solver.  No counterpart in
 * the reference (UMFPACK solves directly). */
int sdddst_solver_history(sdddst_ctx *ctx, double *out, int max_records, int *count, int *strong);
/* Largest N for which an FGMRES solve that does not converge falls back to
 * the dense direct solver (8 N^2 bytes of device memory); above it such a
 * solve returns SDDDST_ERR_NOT_CONVERGED (sdddstCore::Simulation then
 * rejects the attempt and divides the step size by 8).  Default 32768. */
int sdddst_solver_set_direct_limit(sdddst_ctx *ctx, int max_n);
/* With SDDDST_PRECOND_LEARNED: the most columns whose scaled diagonal
 * differs from 1 by more than 0.25 a Newton matrix may have for the
 * network to precondition it; above that (a step size large for the
 * configuration, where the pair blocks alone converge faster) the factor
 * keeps the pair blocks only.  -1: the network always; -2 (default):
 * max(16, N / 256).  The solver history records the preconditioner that
 * served each solve.  No counterpart in the reference. */
int sdddst_set_learned_gate(sdddst_ctx *ctx, int max_strong_rows);

/* ---- Learned preconditioner (SDDDST_PRECOND_LEARNED) ------------------
 * No counterpart in the reference, which solves A delta = g with UMFPACK
 * and no preconditioner (simulation.cpp:386-396).  M = B^-1 + C: B^-1 the
 * pair-block inverse, C on the within-cutoff core of A's pattern with
 * values from a two-layer message-passing network (hidden width 32, FP32
 * weights; sdddst_spai_param_count floats).
 *
 * set/get_weights: install or read the network (needed before a factor
 *   with the learned preconditioner).
 * train_init: weights from the seeded initialiser, Adam state cleared.
 * train_step: on the current Jacobian, K = probes Rademacher probes from
 *   seed: loss L = mean_i log(1 + q_i/eps), q_i = mean_k (A M w_k - w_k)_i^2,
 *   its gradient (grad: sdddst_spai_param_count doubles, nullable) and,
 *   when lr > 0, one Adam step (beta 0.9 / 0.999) of learning rate lr.
 * graph: two-phase download of the last evaluated core graph (Cp[n+1],
 *   Ci[ne]), C's values (Ce[ne], Cd[n]), the network outputs gamma_e[ne],
 *   gamma_d[n] and the pair-block inverse (bd, bo, bcol); any pointer may be
 *   NULL, n_edges receives ne. */
int sdddst_spai_param_count(void);
int sdddst_spai_set_weights(sdddst_ctx *ctx, const float *w, int count);
int sdddst_spai_get_weights(sdddst_ctx *ctx, float *w, int count);
int sdddst_spai_train_init(sdddst_ctx *ctx, uint64_t seed);
int sdddst_spai_train_step(sdddst_ctx *ctx, int probes, uint64_t seed, double eps, double lr, double *loss,
                           double *grad);
int sdddst_spai_graph(sdddst_ctx *ctx, int64_t *n_edges, int *Cp, int *Ci, double *Ce, double *Cd, float *gamma_e,
                      float *gamma_d, double *bd, double *bo, int *bcol);

/* Newton update x[slot] -= delta (simulation.cpp:78-81); delta_host NULL
 * uses the device delta of the last solve. */
int sdddst_newton_update(sdddst_ctx *ctx, int slot, const double *delta_host);

/* Simulation::calculateXError + PrecisionHandler::updateError /
 * getNewStepSize / reset + the max-stepsize clamp (simulation.cpp:398-405,
 * 516, 643-649; precision_handler.cpp:62-119), on the device:
 * max_ratio = max_i |x_big - x_small2|^2 / tol_i, accept = max_ratio < 1,
 * h_next = new step size (clamped to hmax when hmax >= 0; pass a negative
 * hmax for no limit), tolerance reset. */
int sdddst_error_control(sdddst_ctx *ctx, double h, double hmax, int *accept, double *h_next, double *max_ratio,
                         int *selected_id);

/* Reductions of the log (simulation.cpp:460-461, 551, 557, utility.h:92-100):
 * sum |v|, sum v^2 of a speed buffer (sequential order), strain increment
 * sum b*(x[new]-x[old]), order parameter sum b*v. */
int sdddst_speed_sums(sdddst_ctx *ctx, int buf, double *sum_abs, double *sum_sq, double *sum_bv);
int sdddst_strain_increment(sdddst_ctx *ctx, int old_slot, int new_slot, double *out);

/* The end of stage III (simulation.cpp:505-530) in one device round trip:
 * vsq_init / vsq_init2 = sum v^2 of SDDDST_V_INIT / SDDDST_V_INIT2 (the
 * energy terms), the error estimate and controller exactly as
 * sdddst_error_control, and, with strain != 0, the strain increments of the
 * dislocations slot -> first small step -> second small step (as two
 * sdddst_strain_increment calls; the caller adds them only for an accepted
 * step, as the reference does).  The same kernels and values as the
 * separate calls, one host wait instead of five. */
int sdddst_stage_end(sdddst_ctx *ctx, double h, double hmax, int strain, double *vsq_init, double *vsq_init2,
                     int *accept, double *h_next, double *max_ratio, int *selected, double *strain_first,
                     double *strain_second);

int sdddst_synchronize(sdddst_ctx *ctx);

/* AnalyticField::xy / xy_diff_x / xy_diff_x2 (AnalyticField.cpp:145-202,
 * Field.h:34-36) over arrays, on the device: which = 0 xy, 1 xy_diff_x,
 * 3 xy_diff_x2 (the reference's expression trees); which = 2 evaluates the
 * production form of xy.  Backs the AnalyticField class methods and compares
 * per-pair values with the reference's Field virtuals. */
int sdddst_field_eval(sdddst_ctx *ctx, int which, int64_t n, const double *dx, const double *dy, double *out);

/* Device elementary functions (cos, sin, exp, expm1, cosh, sinh, sqrt,
 * pow(x, 1/3) for fn = 0..7) over an array: the bit-compatibility check of
 * the device math against the host C library. */
int sdddst_libm_eval(sdddst_ctx *ctx, int fn, int64_t n, const double *x, double *out);

/* EV analysis of the positions in cfg_slot (the point defects uploaded to the
 * context included): replaces SpectralDecompositor::decompose
 * (spectral_decompositor.cpp:63-201 -- the dense Hessian b_i b_j dtau/dx with
 * its diagonal, then LAPACKE_dsyevd) and the per-mode columns of
 * TimeSeriesProcessor::run (time_series_processor.cpp:40-58, 75-82).
 * eigenvalues[n] ascending; eigenvectors[n * n] row i = eigenvector i (the
 * reference's getEigenVectorElement(i, j)); pn[i] = 1 / sum_j v_ij^4,
 * sum_b[i] = sum_j v_ij b_j, gan[i] = sum_b[i]^2 / eigenvalues[i].  Any
 * output may be NULL.  Needs n * n * 8 bytes of device memory. */
int sdddst_ev_decompose(sdddst_ctx *ctx, int cfg_slot, double *eigenvalues, double *eigenvectors, double *pn,
                        double *gan, double *sum_b);

/* sdddst_ev_decompose plus the eigenvalue derivatives of the reference's
 * --calculate-ev-deriv option (time_series_processor.cpp:92-111 with
 * TimeSeriesProcessor::calculateHessianDerivative, :205-246): given the
 * speeds[n] of the configuration (host array) and the fixed-rate protocol's
 * stress_rate sigma', deriv[m] = sum_ij v_mi v_mj sum_k speeds_k T_ijk /
 * sigma' over the reference's third-derivative tensor T (built on
 * AnalyticField::xy_diff_x2, AnalyticField.cpp:183-202).  speeds and deriv
 * are both given or both NULL (then this is sdddst_ev_decompose).  Needs
 * 3 n * n * 8 bytes of device memory with the derivatives. */
int sdddst_ev_analysis(sdddst_ctx *ctx, int cfg_slot, const double *speeds, double stress_rate, double *eigenvalues,
                       double *eigenvectors, double *pn, double *gan, double *sum_b, double *deriv);

/* Kernel timing with HIP events on the context stream (no host syncs while
 * recording).  enable = 1 clears and starts recording every timer, 2 only the
 * pair kernel's, on one launch in four (each recorded event pair leaves the
 * stream idle for a few microseconds), 0 stops.  kernel is
 * "pair" (the pair-speed tile kernel), "speeds" (one whole speed evaluation)
 * or "jacobian" (one whole assembly); count and total_ms cover the launches
 * recorded since the last enable. */
int sdddst_timing(sdddst_ctx *ctx, int enable);
int sdddst_kernel_stats(sdddst_ctx *ctx, const char *kernel, int *count, double *total_ms);

/* ---- Batched ensembles ------------------------------------------------
 * No counterpart in the reference, which runs one realization per process.
 * num_systems independent realizations (n[s] dislocations each; x, y, b
 * concatenated in realization order) under a constant external stress sigma
 * (0: zero stress), each advanced by Simulation::step (simulation.cpp:
 * 435-650) with the cutoff rc_s = m / sqrt(n[s]), precision and field mode
 * of params (FAST or EXACT), from the step size initial_step.
 *
 * run: attempts of every realization in lock step until each has
 *   target_steps successful steps (or max_attempts ran); one attempt is a
 *   fixed launch sequence, replayed from a captured HIP graph when
 *   use_graph; attempts receives the number run.
 * autotune: times the speeds kernels (symmetric 64x64 tiles, or one lane
 *   per target in workgroups of 64/128/256) and the solver shapes
 *   (256/512/1024 lanes per realization x 4/8/16/64 lanes per matrix row)
 *   on this ensemble, keeps the fastest, writes a report; set_tiles sets
 *   them (speeds_block 0: the symmetric tiles).
 * state: per realization time, step size, successful and failed steps.
 * positions: x of every realization (concatenated).
 * stats: attempts run, FGMRES iterations and solves, dense-LU fallbacks,
 *   capacity regrowths, chosen workgroup sizes (any pointer may be NULL). */
typedef struct sdddst_ensemble sdddst_ensemble;
int sdddst_ensemble_create(int num_systems, const int *n, const double *x, const double *y, const double *b,
                           const sdddst_params *params, double sigma, double initial_step, sdddst_ensemble **out);
int sdddst_ensemble_destroy(sdddst_ensemble *ens);
int sdddst_ensemble_set_tiles(sdddst_ensemble *ens, int speeds_block, int solver_block, int spmv_lanes);
int sdddst_ensemble_autotune(sdddst_ensemble *ens, char *report, int report_len);
int sdddst_ensemble_run(sdddst_ensemble *ens, int64_t target_steps, int64_t max_attempts, int use_graph,
                        int64_t *attempts);
int sdddst_ensemble_state(sdddst_ensemble *ens, double *t, double *h, int64_t *steps, int64_t *failed);
int sdddst_ensemble_positions(sdddst_ensemble *ens, double *x);
int sdddst_ensemble_stats(sdddst_ensemble *ens, int64_t *attempts, int64_t *solver_iters, int64_t *solves,
                          int64_t *lu_fallbacks, int64_t *regrows, int *speeds_block, int *solver_block,
                          int *spmv_lanes);

#ifdef __cplusplus
}
#endif

#endif /* SDDDST_AMD_H */
