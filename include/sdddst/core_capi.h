/*
 * C entry points of the C++ simulator API (sdddstCore::Simulation and
 * SimulationData, include/sdddst/simulation.h) for bindings in other
 * languages: the role the reference's Boost.Python module PySdddstCore plays
 * (bindings.cpp:140-206: SimulationData properties, Simulation.create / run
 * / step / get_time).  Implemented in libsdddstCore.so on top of the C ABI
 * of include/sdddst_amd.h.  Status codes and sdddst_last_error() as there.
 */
#ifndef SDDDST_CORE_CAPI_H
#define SDDDST_CORE_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SDDDST_PROTOCOL_ZERO 0
#define SDDDST_PROTOCOL_FIXED_RATE 1
#define SDDDST_PROTOCOL_SPRING 2
#define SDDDST_PROTOCOL_CONSTANT 3

#define SDDDST_LINSOLVE_AUTO 0
#define SDDDST_LINSOLVE_UMFPACK 1
#define SDDDST_LINSOLVE_DENSE_LU 2
#define SDDDST_LINSOLVE_FGMRES 3

typedef struct sdddst_sim sdddst_sim;

/* SimulationData settings (simulation_data.h:84-254, project_parser.cpp:137-299). */
typedef struct sdddst_sim_options {
    double cutoff_multiplier;  /* cutOffMultiplier                         */
    double precision;          /* prec                                     */
    double initial_step;       /* stepSize at start                        */
    double max_step;           /* maxStepSizeLimit; < 0: no limit          */
    int protocol;              /* SDDDST_PROTOCOL_*                        */
    double stress_param1;      /* rate (fixed-rate, spring) or the stress  */
    double stress_param2;      /* spring constant                          */
    int calculate_strain;      /* calculateStrainDuringSimulation          */
    int order_parameter;       /* orderParameterCalculationIsOn            */
    int linear_solver;         /* SDDDST_LINSOLVE_*                        */
    double solver_tol;         /* FGMRES relative residual                 */
    int solver_max_iter;
    int solver_restart;
    int field_mode;            /* SDDDST_FIELD_* of sdddst_amd.h          */
    int preconditioner;        /* SDDDST_PRECOND_* of sdddst_amd.h        */
    double A, KASQR;           /* point-defect constants (as used)         */
    const char *log_path;      /* 12-column log file; NULL: none           */
    const char *spai_weights;  /* learned-preconditioner weights (raw
                                  float32); NULL: data/spai_gnn_h32.f32
                                  beside libsdddstCore.so                  */
} sdddst_sim_options;

/* Scalars of the state after the last step. */
typedef struct sdddst_sim_state {
    double sim_time, step_size, strain, energy, max_error_ratio_sqr, sum_avg_speed, cutoff;
    long long succesful_steps, failed_steps;
    int last_step_succeeded;
    double last_step_wall_s;
    int solver_last_iters, solver_fallbacks;
    double solver_last_rel_residual;
    long long solver_total_iters, solver_total_solves;
    int solver_precond_fallbacks;  /* learned solves retried with pair blocks */
    long long aborted_attempts;    /* attempts rejected for an unconverged solve */
    int solver_strong_retries;     /* solves retried with the strong-row block */
} sdddst_sim_state;

void sdddst_sim_default_options(sdddst_sim_options *o);
int sdddst_sim_create(int n, const double *x, const double *y, const double *b, int pc, const double *px,
                      const double *py, const sdddst_sim_options *o, sdddst_sim **out);
int sdddst_sim_create_from_files(const char *dconf, const char *pdconf, const sdddst_sim_options *o,
                                 sdddst_sim **out);
int sdddst_sim_destroy(sdddst_sim *s);
/* Simulation::step (simulation.cpp:435-441). */
int sdddst_sim_step(sdddst_sim *s, int *success);
/* Simulation::run with a step-count limit of `successful` more steps. */
int sdddst_sim_run(sdddst_sim *s, int successful);
int sdddst_sim_set_step_size(sdddst_sim *s, double h);
int sdddst_sim_get_state(sdddst_sim *s, sdddst_sim_state *out);
int sdddst_sim_positions(sdddst_sim *s, double *x, int n);
int sdddst_sim_flush_log(sdddst_sim *s);
int sdddst_sim_write_dconf(sdddst_sim *s, const char *path);
const char *sdddst_sim_solver_name(sdddst_sim *s);
const char *sdddst_core_last_error(void);
/* The Field virtuals of AnalyticField (Field.h:34-36, AnalyticField.cpp:
 * 145-202) per point, through the class: which = 0 xy, 1 xy_diff_x,
 * 2 xy_diff_x2 (the bindings' Field.xy / Field.xy_diff_x, bindings.cpp). */
int sdddst_core_field_eval(int which, int n, const double *dx, const double *dy, double *out);
/* The device context the simulation runs on (for sdddst_timing etc.). */
struct sdddst_ctx *sdddst_sim_device_context(sdddst_sim *s);
/* Simulation::run with the limits the SimulationData holds (the bindings'
 * Simulation.run, bindings.cpp:202). */
int sdddst_sim_run_configured(sdddst_sim *s);

/* ---- SimulationData handles: the bindings' SimulationData /
 * SimulationDataObject (bindings.cpp:145-200) ------------------------------
 * A handle owns a shared SimulationData; Simulations created from it share
 * it, so settings changed through the handle act on the next step (the
 * device-side parameters are carried over before each step) and the state
 * (time, steps, positions) reads back live.  Properties are addressed by the
 * bindings' names (cutoff_multiplier, step_size, simulation_time,
 * successful_steps, avalanche_count, ...); booleans read as 0/1.  Array
 * getters return the element count (copying at most cap), <0 on error. */
typedef struct sdddst_data sdddst_data;
/* SimulationData(dislocation file, point-defect file) (simulation_data.cpp:
 * 43-117); either path may be empty.  The field is the analytic one and the
 * stress protocol zero until set. */
int sdddst_data_create(const char *dconf, const char *pdconf, sdddst_data **out);
int sdddst_data_destroy(sdddst_data *d);
int sdddst_data_get(sdddst_data *d, const char *name, double *value);
int sdddst_data_set(sdddst_data *d, const char *name, double value);
/* final_dislocation_configuration_path, sub_config_path */
int sdddst_data_get_string(sdddst_data *d, const char *name, char *buf, int cap);
int sdddst_data_set_string(sdddst_data *d, const char *name, const char *value);
/* which: dislocations, big_step, first_small_step, second_small_step */
int sdddst_data_get_dislocations(sdddst_data *d, const char *which, double *x, double *y, double *b, int cap);
int sdddst_data_set_dislocations(sdddst_data *d, const double *x, const double *y, const double *b, int n);
int sdddst_data_get_point_defects(sdddst_data *d, double *x, double *y, int cap);
int sdddst_data_set_point_defects(sdddst_data *d, const double *x, const double *y, int n);
/* which: g_vec, init_speed, init_speed_2, d_vec */
int sdddst_data_get_vector(sdddst_data *d, const char *which, double *out, int cap);
int sdddst_data_set_vector(sdddst_data *d, const char *which, const double *in, int n);
/* op: 0 readDislocationDataFromFile, 1 writeDislocationDataToFile,
 * 2 readPointDefectDataFromFile, 3 writePointDefectDataToFile */
int sdddst_data_file(sdddst_data *d, int op, const char *path);
int sdddst_data_init_simulation_variables(sdddst_data *d);
int sdddst_data_update_cutoff(sdddst_data *d);
/* tau: kind 0 the Field base (no device path), 1 AnalyticField */
int sdddst_data_set_field(sdddst_data *d, int kind);
/* The data's field per point: which = 0 xy, 1 xy_diff_x, 2 xy_diff_x2. */
int sdddst_data_field_eval(sdddst_data *d, int which, int n, const double *dx, const double *dy, double *out);
/* external_stress: SDDDST_PROTOCOL_* with rate / stress (p1) and spring constant (p2). */
int sdddst_data_set_protocol(sdddst_data *d, int kind, double p1, double p2);
/* op 0: getStress(type); op 1: calculateStress(t, dislocations, type) then
 * getStress(type) (stress_protocol.h:55-81); type: StressProtocolStepType. */
int sdddst_data_protocol(sdddst_data *d, int op, double t, int type, double *out);
/* Simulation::create (simulation.cpp:658-668): a Simulation on the handle's
 * SimulationData (shared). */
int sdddst_sim_create_from_data(sdddst_data *d, sdddst_sim **out);

/* PrecisionHandler (precision_handler.h:28-60, the bindings' class of the
 * same name) on the host.  op: 0 setSize(a), 1 getSize, 2 reset,
 * 3 updateTolerance(a = distance^2, id), 4 updateError(a = error, id),
 * 5 getNewStepSize(a = old step), 6 getMinPrecisity, 7 setMinPrecisity(a),
 * 8 getMaxErrorRatioSqr; *out (nullable) receives the result. */
typedef struct sdddst_ph sdddst_ph;
int sdddst_ph_create(sdddst_ph **out);
int sdddst_ph_destroy(sdddst_ph *p);
int sdddst_ph_call(sdddst_ph *p, int op, double a, unsigned int id, double *out);

/* utility.h (the bindings' module functions): which 0 normalize(a),
 * 1 X(a), 2 X2(a), 3 E(a, b, k), 4 X_dx(a), 5 X2_dx(a), 6 E_dx(a, b, k),
 * 7 get_wall_time(), 8 get_cpu_time(). */
double sdddst_core_utility(int which, double a, double b, double k);

#ifdef __cplusplus
}
#endif

#endif
