// Internal state of one simulation context: device buffers, the stream and
// the launch helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sdddst_amd.h"

namespace sdamd {

void set_error(const char *fmt, ...);

#define SD_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            ::sdamd::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                               __LINE__);                                                       \
            return SDDDST_ERR_HIP;                                                              \
        }                                                                                       \
    } while (0)

#define SD_CHECK(cond, code, ...)            \
    do {                                     \
        if (!(cond)) {                       \
            ::sdamd::set_error(__VA_ARGS__); \
            return (code);                   \
        }                                    \
    } while (0)

struct DenseLu;
struct Fgmres;

// Device scalars of the step controller (one per context).
struct DevScalars {
    double max_ratio;      // PrecisionHandler::maxErrorRatioSqr
    double h_next;         // getNewStepSize result (after the clamp)
    int accept;            // maxErrorRatioSqr < 1
    int selected;          // PrecisionHandler::selectedID
    double sum_abs, sum_sq, sum_bv, strain;
};

// Completion signal of the host waits: a one-lane kernel copies a few small
// device results into this block of host-coherent mapped memory and then
// publishes a sequence number with a system-scope release; the host polls
// the number instead of blocking in the runtime.
struct HostSignal {
    uint64_t seq;
    uint64_t pad;
    uint64_t payload[14];
};
static_assert(sizeof(DevScalars) % 8 == 0 && sizeof(DevScalars) <= 14 * 8, "DevScalars fits the signal payload");

constexpr int kTimePair = 1, kTimeSpeeds = 2, kTimeJacobian = 4;

// A column whose scaled diagonal differs from 1 by more than this is
// "strongly shifted": the Newton matrix is then far from the identity there
// (a close pair or a pinned dislocation at a large step size).
constexpr double kStrongDiag = 0.25;

struct Context {
    int n = 0, pc = 0;
    sdddst_params p{};
    double cutoff = 0, cutoff_sqr = 0, one_per_cutoff_sqr = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    // positions: x per slot, y and b shared
    double *x[SDDDST_NUM_SLOTS] = {};
    double *y = nullptr, *b = nullptr;
    double *px = nullptr, *py = nullptr;
    // speeds
    double *v[SDDDST_NUM_VBUFS] = {};
    double *tol = nullptr;              // tolerance per dislocation
    unsigned long long *minr2 = nullptr; // per-call minimum r^2 bits
    double *partial = nullptr;          // [nb][n] pair partial sums
    int nb = 0;                         // number of 64-wide blocks
    // pair-tile order: perm[p] = dislocation at tile position p (curve order)
    int *perm = nullptr, *xkey = nullptr, *xkey_sorted = nullptr, *xidx = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int order_age = 0;                  // speeds calls since the order was last computed
    double *ytc = nullptr, *yts = nullptr;   // cos, sin of 2 pi y per dislocation (set with y)
    double *yw = nullptr;                    // y - rint(y) per dislocation (set with y)
    double *xw = nullptr, *xep = nullptr, *xem = nullptr;   // x - rint(x), e^{+-2 pi (x - rint(x))}
    double *bext = nullptr;                  // per 64-block extents of xw and yw (pairtile::kExtWords each)
    // series of the x-separated tiles: source moments per block, their
    // n = 1 squares, and the tiles' share of each speed (sorted positions)
    double *farx_mom = nullptr, *farx_mom1 = nullptr, *vfar = nullptr;
    // Newton
    double *dvec = nullptr, *g = nullptr, *delta = nullptr;
    // ||g||^2 formed by the residual kernel (block partials, last-block
    // ticket), valid until a Krylov solve takes it
    double *gnorm2 = nullptr, *gpart = nullptr;
    unsigned *gticket = nullptr;
    bool gnorm2_valid = false;
    // Jacobian (CSC)
    int *Ap = nullptr, *Ai = nullptr, *diag = nullptr, *colcount = nullptr;
    // the row indices again as 16 bits (N <= 65536), for the Krylov products:
    // 10 instead of 12 bytes per entry; valid for the pattern of the last fill
    unsigned short *Ai16 = nullptr;
    bool ai16_valid = false;
    double *Ax = nullptr;               // final values (formed on request: ensure_ax)
    // Ax holds the current matrix; else it is (Araw * ax_factor) times the
    // row scales of colinfo, +1 on the diagonal
    bool ax_valid = false;
    double ax_factor = 1.0;
    double *AxRow = nullptr;            // CSR-of-A values (same index arrays): for SpMV
    int64_t nnz = 0, cap = 0;
    double *jtau = nullptr, *jmult = nullptr;   // tau_x and taper per strictly-lower entry
    int *jmirror = nullptr;                     // position of the mirror entry
    bool aux_valid = false;                     // jtau / jmult / jmirror hold the current pattern's
    // x at the last pattern build (written by its values pass) and the
    // squared cutoff radius then: an entry's pair lies within the cutoff
    // (untapered) where its r^2 from these positions is not above jrc2
    double *jx = nullptr;
    double jrc2 = 0;
    // the same relation as a [n][ceil(n/64)] bitmask, written by the count
    // pass of a build while the learned preconditioner is selected
    unsigned long long *jcmask = nullptr;
    int *jccnt = nullptr;                       // set bits per column of jcmask
    bool jcmask_valid = false;
    double *jpd = nullptr;                      // point-defect sum per column (h-free)
    unsigned long long *jmask = nullptr;        // [n][ceil(n/64)] membership bitmask
    int *jwstart = nullptr;                     // [n][ceil(n/64)] position of each mask word's first row
    int *cell = nullptr, *csorted = nullptr;    // cell of each dislocation, dislocations by cell
    double *csx = nullptr, *csy = nullptr;      // their positions in that order
    int *ccount = nullptr, *cstart = nullptr, *ccursor = nullptr;
    int cell_cap = 0;
    bool pattern_valid = false;
    int pattern_slot = -1;                      // slot whose positions built the pattern
    bool jac_valid = false;
    bool jac_symmetric = true;
    bool direct_once = false;                   // this Jacobian goes to the direct solver
    double jac_h = 0;
    // fused assembly: unscaled values (Araw, nnz), their column sums after
    // pass 2 of the reference (sraw, n) at step size raw_h, and the pair
    // partners of the scaled matrix written by the scaling pass
    double *Araw = nullptr, *sraw = nullptr;
    // per column j, written with d_j: {(1 + d_j) / 2, |s_j (1 + d_j) / 2 + 1|}
    // (the row scale and the scaled diagonal's magnitude, as the scaling
    // pass forms them)
    double2 *colinfo = nullptr;
    bool raw_valid = false;
    double raw_h = 0;
    // step size of the off-diagonal unsigned values in Araw (written by the
    // values pass of a pattern build, or by k_jac_raw for another h)
    double araw_h = __builtin_nan("");   // NaN: none
    int *partner = nullptr;
    double *strength = nullptr;
    bool partner_valid = false;
    int *d_flags = nullptr;             // device flags: [0] zero-valued upper entries found,
                                        // [2..3] entry count, [4] strongly shifted diagonals
    // columns with |A_jj - 1| > kStrongDiag in the last full assembly, and its h
    int diag_strong = 0;
    double diag_strong_h = 0.0;
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    // controller scalars
    DevScalars *ds = nullptr;           // kNumDevScalars records; [0] for the single-result calls
    DevScalars *hs = nullptr;           // pinned host mirror
    HostSignal *sig = nullptr;          // host view of the completion signal (mapped, coherent)
    HostSignal *sig_dev = nullptr;      // its device address
    uint64_t sig_seq = 0;
    // solver
    int solver = SDDDST_SOLVER_DENSE_LU;
    double solver_tol = 1e-10;
    int solver_max_iter = 200, solver_restart = 30;
    int precond = SDDDST_PRECOND_PAIR;
    // the preconditioner the current factorisation serves: precond, except
    // that the learned one gives way to the pair blocks on a matrix with more
    // strongly shifted diagonals than learned_gate (-1: never; -2: the
    // default max(16, N / 256))
    int precond_active = SDDDST_PRECOND_PAIR;
    int learned_gate = -2;
    int precond_override = -1;          // >= 0: preconditioner of the current retry
    int precond_fallbacks = 0;          // learned solves retried with the pair blocks
    struct Spai *spai = nullptr;       // learned preconditioner state
    DenseLu *lu = nullptr;
    Fgmres *gm = nullptr;
    int last_iters = 0, fallbacks = 0;
    int64_t total_iters = 0, total_solves = 0;
    double last_rel_res = 0, last_gnorm = 0;
    int strong_solves = 0;              // solves retried with the strong-row block preconditioner
    int direct_max_n = 32768;           // largest N the dense direct fallback may factor (8 N^2 bytes)
    // the most recent iterative-solver attempts, oldest first once full:
    // {preconditioner, iterations, relative residual, ||g||, outcome, h,
    // rows of the strong-row block} with outcome 0 converged, 1 not
    // converged, 2 solved by the direct solver after it
    static constexpr int kSolveLogMax = 1024, kSolveLogFields = 7;
    double solve_log[kSolveLogMax][kSolveLogFields];
    int solve_log_len = 0, solve_log_next = 0;
    void log_solve(int pc, int iters, double rel, double gnorm, int outcome, int strong_rows = 0)
    {
        double *r = solve_log[solve_log_next];
        r[0] = pc;
        r[1] = iters;
        r[2] = rel;
        r[3] = gnorm;
        r[4] = outcome;
        r[5] = jac_h;
        r[6] = strong_rows;
        solve_log_next = (solve_log_next + 1) % kSolveLogMax;
        if (solve_log_len < kSolveLogMax) ++solve_log_len;
    }
    // timing: recorded event pairs per kernel kind, read back on request
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // timers recording: kTimePair | kTimeSpeeds | kTimeJacobian (each event
    // pair also leaves the stream idle for a few microseconds)
    int timing = 0;
    struct Timer {
        std::vector<hipEvent_t> begin, end;
        int used = 0;
        double total_ms = 0;
        int count = 0;
        int every = 1;          // record one call in `every`
        long calls = 0;
        bool active = false;    // this call is recorded
    } t_pair, t_speeds, t_jac;
};

// Record the start / end of a timed region on the context stream.
int timer_begin(Context &c, Context::Timer &t);
int timer_end(Context &c, Context::Timer &t);
int timer_collect(Context &c, Context::Timer &t);

// cutoff of SimulationData::updateCutOff (simulation_data.cpp:208-218)
void compute_cutoff(Context &c);

// kernel launchers (return SDDDST status)
int launch_speeds(Context &c, int slot, int out_buf, double sigma, int update_tol);
int pair_order_alloc(Context &c);   // buffers of the x ordering of the pair tiles
int launch_ytrig(Context &c);       // cos / sin of 2 pi y after an upload of y
// The EV analysis of the positions in slot (ev_analysis.hip); outputs nullable.
// With speeds (host) and deriv both given, also the eigenvalue derivatives
// deriv[m] = v_m^T C v_m / rate of the reference's fixed-rate option.
int ev_decompose(Context &c, int slot, double *eigenvalues, double *eigenvectors, double *pn, double *gan,
                 double *sum_b, const double *speeds = nullptr, double rate = 1.0, double *deriv = nullptr);
// Waits until the work enqueued on the stream so far is done, copying `words`
// 64-bit words from device address src (if any) to dst on the way: a signal
// kernel publishes them in HostSignal and the host polls it (a bounded spin,
// then hipStreamSynchronize, which also reports device errors).
int host_wait(Context &c, const void *src = nullptr, int words = 0, void *dst = nullptr);
int launch_jacobian(Context &c, int slot, double h, bool reuse_pattern);
// The column-major values Ax of the current matrix (the assembly writes the
// row-major AxRow only).
int ensure_ax(Context &c);
int launch_residual(Context &c, double h, int new_slot, int old_slot, int csp, int isp);
int launch_newton_update(Context &c, int slot);
int launch_error_control(Context &c, double h, double hmax);
// the sums / the increment into dst (nullptr: c.ds)
int launch_speed_sums(Context &c, int buf, DevScalars *dst = nullptr);
int launch_strain(Context &c, int old_slot, int new_slot, DevScalars *dst = nullptr);
// the end of stage III in one record (c.ds[kStageOut]): see sdddst_stage_end
constexpr int kStageSumA = 1, kStageSumB = 2, kStageStrainA = 3, kStageStrainB = 4, kStageOut = 5, kNumDevScalars = 6;
int launch_stage_pack(Context &c);
int launch_reset_tolerance(Context &c);
int field_eval(Context &c, int which, int64_t n, const double *dx, const double *dy, double *out);
int libm_eval(Context &c, int fn, int64_t n, const double *x, double *out);
int launch_scan(Context &c, int n, const int *cnt, int *out, long long *d_total);
int solver_factor(Context &c);
int solver_solve(Context &c);
void solver_free(Context &c);
void spai_free(Context &c);

}  // namespace sdamd

struct sdddst_ctx {
    sdamd::Context c;
};
