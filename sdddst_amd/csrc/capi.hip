// The C ABI of include/sdddst_amd.h: argument checking, device memory and
// stream ownership, and the thin calls into the kernel launchers.
#include <hip/hip_runtime.h>

#include <cstring>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>

#include "context.h"
#include "solvers.h"

namespace sdamd {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// simulation_data.cpp:208-218
void compute_cutoff(Context &c)
{
    double multiplier = c.p.cutoff_multiplier;
    if (c.p.cutoff_multiplier >= 1. / 12. * sqrt(2.0 * double(c.n))) multiplier = 1e20;
    c.cutoff = 1. / sqrt((double)c.n) * multiplier;
    c.cutoff_sqr = c.cutoff * c.cutoff;
    c.one_per_cutoff_sqr = 1. / c.cutoff_sqr;
}

static int check_params(const sdddst_params *p)
{
    SD_CHECK(p, SDDDST_ERR_ARG, "params is NULL");
    SD_CHECK(p->precision > 0, SDDDST_ERR_ARG, "precision must be > 0");
    SD_CHECK(p->iteration_count >= 1, SDDDST_ERR_ARG, "iteration_count must be >= 1");
    SD_CHECK(p->field_mode >= SDDDST_FIELD_FAST && p->field_mode <= SDDDST_FIELD_ORDERED, SDDDST_ERR_ARG,
             "unknown field mode %d", p->field_mode);
    return SDDDST_OK;
}

static void free_all(Context &c)
{
    solver_free(c);
    spai_free(c);
    for (int s = 0; s < SDDDST_NUM_SLOTS; ++s)
        if (c.x[s]) (void)hipFree(c.x[s]);
    for (int s = 0; s < SDDDST_NUM_VBUFS; ++s)
        if (c.v[s]) (void)hipFree(c.v[s]);
    void *ptrs[] = {c.y, c.b, c.px, c.py, c.tol, c.minr2, c.partial, c.dvec, c.g, c.delta, c.Ap, c.Ai, c.diag,
                    c.colcount, c.Ax, c.AxRow, c.d_flags, c.ds, c.jtau, c.jmult, c.jmirror, c.jmask, c.jpd, c.jx,
                    c.jcmask, c.jccnt,
                    c.cell, c.csorted, c.ccount, c.cstart, c.ccursor, c.perm, c.xkey, c.xkey_sorted, c.xidx,
                    c.sort_tmp, c.Araw, c.sraw, c.partner, c.strength, c.jwstart, c.csx, c.csy, c.colinfo,
                    c.ytc, c.yts, c.yw, c.xw, c.xep, c.xem, c.bext, c.farx_mom, c.farx_mom1, c.vfar, c.Ai16, c.gnorm2, c.gpart, c.gticket};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c.hs) (void)hipHostFree(c.hs);
    if (c.sig) (void)hipHostFree(c.sig);
    for (Context::Timer *t : {&c.t_pair, &c.t_speeds, &c.t_jac}) {
        for (hipEvent_t e : t->begin) (void)hipEventDestroy(e);
        for (hipEvent_t e : t->end) (void)hipEventDestroy(e);
    }
    if (c.ev0) (void)hipEventDestroy(c.ev0);
    if (c.ev1) (void)hipEventDestroy(c.ev1);
    if (c.stream) (void)hipStreamDestroy(c.stream);
}

template <class T>
static int dalloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc of %zu bytes failed", sizeof(T) * count);
        return SDDDST_ERR_NOMEM;
    }
    return SDDDST_OK;
}

static int create(int n, int pc, const sdddst_params *params, Context &c)
{
    SD_CHECK(n >= 2, SDDDST_ERR_ARG, "need at least 2 dislocations (n=%d)", n);
    SD_CHECK(pc >= 0, SDDDST_ERR_ARG, "negative point-defect count");
    int rc = check_params(params);
    if (rc) return rc;
    int ndev = 0;
    SD_HIP(hipGetDeviceCount(&ndev));
    SD_CHECK(ndev > 0, SDDDST_ERR_HIP, "no HIP device");
    SD_HIP(hipGetDevice(&c.device));
    // The step waits on the device for small results many times per step
    // (Krylov coefficients, step acceptance): spin on completion instead of
    // yielding the thread, which otherwise adds a wake-up latency to every
    // wait.  Best effort: a runtime already active with other flags keeps them.
    {
        unsigned flags = 0;
        if (hipGetDeviceFlags(&flags) == hipSuccess && (flags & hipDeviceScheduleMask) != hipDeviceScheduleSpin)
            (void)hipSetDeviceFlags((flags & ~hipDeviceScheduleMask) | hipDeviceScheduleSpin);
        (void)hipGetLastError();
    }
    c.n = n;
    c.pc = pc;
    c.p = *params;
    compute_cutoff(c);
    c.nb = (n + 63) / 64;
    SD_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    SD_HIP(hipEventCreate(&c.ev0));
    SD_HIP(hipEventCreate(&c.ev1));
    for (int s = 0; s < SDDDST_NUM_SLOTS; ++s)
        if ((rc = dalloc(&c.x[s], n))) return rc;
    for (int s = 0; s < SDDDST_NUM_VBUFS; ++s)
        if ((rc = dalloc(&c.v[s], n))) return rc;
    if ((rc = dalloc(&c.y, n)) || (rc = dalloc(&c.b, n)) || (rc = dalloc(&c.px, pc)) || (rc = dalloc(&c.py, pc)) ||
        (rc = dalloc(&c.tol, n)) || (rc = dalloc(&c.minr2, n)) || (rc = dalloc(&c.partial, (size_t)c.nb * n)) ||
        (rc = dalloc(&c.dvec, n)) || (rc = dalloc(&c.g, n)) || (rc = dalloc(&c.delta, n)) ||
        (rc = dalloc(&c.Ap, (size_t)n + 1)) || (rc = dalloc(&c.diag, n)) || (rc = dalloc(&c.colcount, n)) ||
        (rc = dalloc(&c.d_flags, 8)) || (rc = dalloc(&c.ds, kNumDevScalars)) || (rc = dalloc(&c.jpd, n)) ||
        (rc = dalloc(&c.sraw, n)) || (rc = dalloc(&c.partner, n)) || (rc = dalloc(&c.strength, n)) ||
        (rc = dalloc(&c.colinfo, n)) || (rc = dalloc(&c.gnorm2, 1)) || (rc = dalloc(&c.gpart, (size_t)c.nb * 4)) ||
        (rc = dalloc(&c.gticket, 1)) || (rc = dalloc(&c.jx, n)) ||
        (rc = pair_order_alloc(c)))
        return rc;
    SD_HIP(hipMemsetAsync(c.gticket, 0, sizeof(unsigned), c.stream));
    SD_HIP(hipHostMalloc((void **)&c.hs, sizeof(DevScalars), hipHostMallocDefault));
    SD_HIP(hipHostMalloc((void **)&c.sig, sizeof(HostSignal), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c.sig, 0, sizeof(HostSignal));
    SD_HIP(hipHostGetDevicePointer((void **)&c.sig_dev, c.sig, 0));
    SD_HIP(hipMemsetAsync(c.dvec, 0, sizeof(double) * n, c.stream));
    SD_HIP(hipMemsetAsync(c.x[0], 0, sizeof(double) * n, c.stream));
    rc = launch_reset_tolerance(c);
    if (rc) return rc;
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

// Event pairs grow in chunks; elapsed times are read only in timer_collect.
int timer_begin(Context &c, Context::Timer &t)
{
    t.active = t.calls++ % t.every == 0;
    if (!t.active) return SDDDST_OK;
    if (t.used == (int)t.begin.size()) {
        if (t.used >= 4096) {
            int rc = timer_collect(c, t);
            if (rc) return rc;
        } else {
            hipEvent_t a, b;
            SD_HIP(hipEventCreate(&a));
            SD_HIP(hipEventCreate(&b));
            t.begin.push_back(a);
            t.end.push_back(b);
        }
    }
    SD_HIP(hipEventRecord(t.begin[t.used], c.stream));
    return SDDDST_OK;
}

int timer_end(Context &c, Context::Timer &t)
{
    if (!t.active) return SDDDST_OK;
    SD_HIP(hipEventRecord(t.end[t.used], c.stream));
    t.used++;
    return SDDDST_OK;
}

int timer_collect(Context &c, Context::Timer &t)
{
    if (t.used == 0) return SDDDST_OK;
    SD_HIP(hipEventSynchronize(t.end[t.used - 1]));
    for (int k = 0; k < t.used; ++k) {
        float ms = 0;
        SD_HIP(hipEventElapsedTime(&ms, t.begin[k], t.end[k]));
        t.total_ms += ms;
        t.count++;
    }
    t.used = 0;
    return SDDDST_OK;
}

static int valid_slot(int s) { return s >= 0 && s < SDDDST_NUM_SLOTS; }
static int valid_buf(int s) { return s >= 0 && s < SDDDST_NUM_VBUFS; }

}  // namespace sdamd

using namespace sdamd;

#define CTX_CHECK(ctx) SD_CHECK((ctx) != nullptr, SDDDST_ERR_ARG, "context is NULL")

extern "C" {

const char *sdddst_version(void) { return "sdddst_amd 0.1 (gfx950, FP64)"; }

const char *sdddst_last_error(void) { return g_last_error.c_str(); }

int sdddst_ctx_create(int n, int pc, const sdddst_params *params, sdddst_ctx **out)
{
    SD_CHECK(out != nullptr, SDDDST_ERR_ARG, "out is NULL");
    *out = nullptr;
    sdddst_ctx *ctx = new (std::nothrow) sdddst_ctx();
    SD_CHECK(ctx != nullptr, SDDDST_ERR_NOMEM, "host allocation failed");
    int rc = create(n, pc, params, ctx->c);
    if (rc) {
        free_all(ctx->c);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return SDDDST_OK;
}

int sdddst_ctx_destroy(sdddst_ctx *ctx)
{
    if (!ctx) return SDDDST_OK;
    (void)hipStreamSynchronize(ctx->c.stream);
    free_all(ctx->c);
    delete ctx;
    return SDDDST_OK;
}

int sdddst_set_params(sdddst_ctx *ctx, const sdddst_params *params)
{
    CTX_CHECK(ctx);
    int rc = check_params(params);
    if (rc) return rc;
    ctx->c.p = *params;
    compute_cutoff(ctx->c);
    ctx->c.jac_valid = false;
    ctx->c.pattern_valid = false;
    return SDDDST_OK;
}

int sdddst_get_cutoff(sdddst_ctx *ctx, double *cutoff, double *cutoff_sqr)
{
    CTX_CHECK(ctx);
    if (cutoff) *cutoff = ctx->c.cutoff;
    if (cutoff_sqr) *cutoff_sqr = ctx->c.cutoff_sqr;
    return SDDDST_OK;
}

int sdddst_upload_state(sdddst_ctx *ctx, int slot, const double *x, const double *y, const double *b, int n)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(slot), SDDDST_ERR_ARG, "bad slot %d", slot);
    SD_CHECK(n == c.n, SDDDST_ERR_ARG, "n=%d does not match the context (%d)", n, c.n);
    SD_CHECK(x != nullptr, SDDDST_ERR_ARG, "x is NULL");
    SD_HIP(hipMemcpyAsync(c.x[slot], x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    if (y) {
        SD_HIP(hipMemcpyAsync(c.y, y, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        const int rc = launch_ytrig(c);
        if (rc) return rc;
    }
    if (b) SD_HIP(hipMemcpyAsync(c.b, b, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    c.order_age = 0;   // new positions: the next speeds call orders them afresh
    // The assembled matrix stays valid.  Reusing its pattern (rebuild = 2) is
    // the caller's assertion about x; new y or b rule it out.
    if (y != nullptr || b != nullptr) c.pattern_valid = false;
    return SDDDST_OK;
}

int sdddst_download_x(sdddst_ctx *ctx, int slot, double *x, int n)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(slot) && x && n == c.n, SDDDST_ERR_ARG, "bad arguments");
    SD_HIP(hipMemcpyAsync(x, c.x[slot], sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_copy_slot(sdddst_ctx *ctx, int dst, int src)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(dst) && valid_slot(src), SDDDST_ERR_ARG, "bad slot");
    if (dst != src)
        SD_HIP(hipMemcpyAsync(c.x[dst], c.x[src], sizeof(double) * c.n, hipMemcpyDeviceToDevice, c.stream));
    return SDDDST_OK;
}

int sdddst_swap_slots(sdddst_ctx *ctx, int a, int b)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(a) && valid_slot(b), SDDDST_ERR_ARG, "bad slot");
    double *t = c.x[a];
    c.x[a] = c.x[b];
    c.x[b] = t;
    return SDDDST_OK;
}

int sdddst_upload_points(sdddst_ctx *ctx, const double *px, const double *py, int pc)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(pc == c.pc, SDDDST_ERR_ARG, "pc=%d does not match the context (%d)", pc, c.pc);
    if (pc == 0) return SDDDST_OK;
    SD_CHECK(px && py, SDDDST_ERR_ARG, "NULL point arrays");
    SD_HIP(hipMemcpyAsync(c.px, px, sizeof(double) * pc, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipMemcpyAsync(c.py, py, sizeof(double) * pc, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_speeds(sdddst_ctx *ctx, int cfg_slot, int out_buf, double sigma, int update_tol, double *v_out)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(cfg_slot) && valid_buf(out_buf), SDDDST_ERR_ARG, "bad slot or buffer");
    int rc = launch_speeds(c, cfg_slot, out_buf, sigma, update_tol);
    if (rc) return rc;
    if (v_out) {
        SD_HIP(hipMemcpyAsync(v_out, c.v[out_buf], sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipStreamSynchronize(c.stream));
    }
    return SDDDST_OK;
}

int sdddst_get_tolerance(sdddst_ctx *ctx, double *tol_out, int n)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(tol_out && n == c.n, SDDDST_ERR_ARG, "bad arguments");
    SD_HIP(hipMemcpyAsync(tol_out, c.tol, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_reset_tolerance(sdddst_ctx *ctx)
{
    CTX_CHECK(ctx);
    return launch_reset_tolerance(ctx->c);
}

int sdddst_download_speed(sdddst_ctx *ctx, int buf, double *v_out, int n)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_buf(buf) && v_out && n == c.n, SDDDST_ERR_ARG, "bad arguments");
    SD_HIP(hipMemcpyAsync(v_out, c.v[buf], sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_upload_speed(sdddst_ctx *ctx, int buf, const double *v, int n)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_buf(buf) && v && n == c.n, SDDDST_ERR_ARG, "bad arguments");
    SD_HIP(hipMemcpyAsync(c.v[buf], v, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_copy_speed(sdddst_ctx *ctx, int dst, int src)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_buf(dst) && valid_buf(src), SDDDST_ERR_ARG, "bad buffer");
    if (dst != src)
        SD_HIP(hipMemcpyAsync(c.v[dst], c.v[src], sizeof(double) * c.n, hipMemcpyDeviceToDevice, c.stream));
    return SDDDST_OK;
}

int sdddst_jacobian_csc(sdddst_ctx *ctx, int cfg_slot, double h, int rebuild, int *Ap, int *Ai, double *Ax,
                        double *dVec, int *diag, int64_t *nnz)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(cfg_slot), SDDDST_ERR_ARG, "bad slot");
    SD_CHECK(rebuild >= 0 && rebuild <= 2, SDDDST_ERR_ARG, "rebuild must be 0, 1 or 2");
    if (rebuild) {
        int rc = launch_jacobian(c, cfg_slot, h, rebuild == 2);
        if (rc) return rc;
    }
    SD_CHECK(c.jac_valid, SDDDST_ERR_STATE, "no Jacobian assembled");
    if (nnz) *nnz = c.nnz;
    if (Ap) {
        SD_CHECK(Ai && Ax, SDDDST_ERR_ARG, "Ai/Ax are NULL");
        int rc = ensure_ax(c);
        if (rc) return rc;
        SD_HIP(hipMemcpyAsync(Ap, c.Ap, sizeof(int) * (c.n + 1), hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipMemcpyAsync(Ai, c.Ai, sizeof(int) * c.nnz, hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipMemcpyAsync(Ax, c.Ax, sizeof(double) * c.nnz, hipMemcpyDeviceToHost, c.stream));
    }
    if (dVec) SD_HIP(hipMemcpyAsync(dVec, c.dvec, sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
    if (diag) SD_HIP(hipMemcpyAsync(diag, c.diag, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
    // wait only for copies to the host: an assembly alone stays queued ahead
    // of the factorisation the caller enqueues next
    if (Ap || dVec || diag) SD_HIP(hipStreamSynchronize(c.stream));
    return SDDDST_OK;
}

int sdddst_residual(sdddst_ctx *ctx, double h, int new_slot, int old_slot, int csp_buf, int isp_buf, double *g_out)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(new_slot) && valid_slot(old_slot) && valid_buf(csp_buf) && valid_buf(isp_buf),
             SDDDST_ERR_ARG, "bad slot or buffer");
    int rc = launch_residual(c, h, new_slot, old_slot, csp_buf, isp_buf);
    if (rc) return rc;
    if (g_out) {
        SD_HIP(hipMemcpyAsync(g_out, c.g, sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipStreamSynchronize(c.stream));
    }
    return SDDDST_OK;
}

int sdddst_solver_select(sdddst_ctx *ctx, int solver, double rel_tol, int max_iter, int restart)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(solver == SDDDST_SOLVER_DENSE_LU || solver == SDDDST_SOLVER_FGMRES || solver == SDDDST_SOLVER_HOST,
             SDDDST_ERR_ARG, "unknown solver %d", solver);
    c.solver = solver;
    if (rel_tol > 0) c.solver_tol = rel_tol;
    if (max_iter > 0) c.solver_max_iter = max_iter;
    if (restart > 0) c.solver_restart = restart;
    return SDDDST_OK;
}

int sdddst_set_preconditioner(sdddst_ctx *ctx, int precond)
{
    CTX_CHECK(ctx);
    SD_CHECK(precond >= SDDDST_PRECOND_JACOBI && precond <= SDDDST_PRECOND_STRONG, SDDDST_ERR_ARG,
             "unknown preconditioner %d", precond);
    ctx->c.precond = precond;
    return SDDDST_OK;
}

int sdddst_solver_factor(sdddst_ctx *ctx)
{
    CTX_CHECK(ctx);
    return solver_factor(ctx->c);
}

int sdddst_solver_solve(sdddst_ctx *ctx, double *delta_out)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    int rc = solver_solve(c);
    if (rc) return rc;
    if (delta_out) {
        SD_HIP(hipMemcpyAsync(delta_out, c.delta, sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
        SD_HIP(hipStreamSynchronize(c.stream));
    }
    return SDDDST_OK;
}

int sdddst_solver_stats(sdddst_ctx *ctx, int *last_iters, double *last_rel_residual, int *fallbacks,
                        int64_t *total_iters, int64_t *total_solves)
{
    CTX_CHECK(ctx);
    if (total_iters) *total_iters = ctx->c.total_iters;
    if (total_solves) *total_solves = ctx->c.total_solves;
    if (last_iters) *last_iters = ctx->c.last_iters;
    if (last_rel_residual) *last_rel_residual = ctx->c.last_rel_res;
    if (fallbacks) *fallbacks = ctx->c.fallbacks;
    return SDDDST_OK;
}

int sdddst_solver_precond_fallbacks(sdddst_ctx *ctx, int *count)
{
    CTX_CHECK(ctx);
    SD_CHECK(count, SDDDST_ERR_ARG, "count is required");
    *count = ctx->c.precond_fallbacks;
    return SDDDST_OK;
}

int sdddst_solver_history(sdddst_ctx *ctx, double *out, int max_records, int *count, int *strong)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(count && max_records >= 0 && (out || max_records == 0), SDDDST_ERR_ARG, "bad arguments");
    const int k = c.solve_log_len < max_records ? c.solve_log_len : max_records;
    // the k most recent records, oldest first
    const int first = (c.solve_log_next - k + Context::kSolveLogMax) % Context::kSolveLogMax;
    for (int r = 0; r < k; ++r) {
        const double *src = c.solve_log[(first + r) % Context::kSolveLogMax];
        for (int f = 0; f < Context::kSolveLogFields; ++f) out[Context::kSolveLogFields * r + f] = src[f];
    }
    *count = k;
    if (strong) *strong = c.strong_solves;
    if (max_records > 0) c.solve_log_len = 0;
    return SDDDST_OK;
}

int sdddst_solver_set_direct_limit(sdddst_ctx *ctx, int max_n)
{
    CTX_CHECK(ctx);
    SD_CHECK(max_n >= 0, SDDDST_ERR_ARG, "negative limit");
    ctx->c.direct_max_n = max_n;
    return SDDDST_OK;
}

int sdddst_set_learned_gate(sdddst_ctx *ctx, int max_strong_rows)
{
    CTX_CHECK(ctx);
    SD_CHECK(max_strong_rows >= -2, SDDDST_ERR_ARG, "learned gate %d below -2", max_strong_rows);
    ctx->c.learned_gate = max_strong_rows;
    return SDDDST_OK;
}

int sdddst_spai_param_count(void) { return spai_param_count(); }

int sdddst_spai_set_weights(sdddst_ctx *ctx, const float *w, int count)
{
    CTX_CHECK(ctx);
    SD_CHECK(w && count == spai_param_count(), SDDDST_ERR_ARG, "expected %d weights, got %d", spai_param_count(),
             count);
    for (int t = 0; t < count; ++t) SD_CHECK(std::isfinite(w[t]), SDDDST_ERR_ARG, "weight %d is not finite", t);
    return spai_set_weights(ctx->c, w);
}

int sdddst_spai_get_weights(sdddst_ctx *ctx, float *w, int count)
{
    CTX_CHECK(ctx);
    SD_CHECK(w && count == spai_param_count(), SDDDST_ERR_ARG, "expected room for %d weights, got %d",
             spai_param_count(), count);
    return spai_get_weights(ctx->c, w);
}

int sdddst_spai_train_init(sdddst_ctx *ctx, uint64_t seed)
{
    CTX_CHECK(ctx);
    return spai_train_init(ctx->c, seed);
}

int sdddst_spai_train_step(sdddst_ctx *ctx, int probes, uint64_t seed, double eps, double lr, double *loss,
                           double *grad)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(probes >= 1 && probes <= 64, SDDDST_ERR_ARG, "probes must be in [1, 64]");
    SD_CHECK(eps > 0 && std::isfinite(eps), SDDDST_ERR_ARG, "eps must be positive");
    SD_CHECK(std::isfinite(lr), SDDDST_ERR_ARG, "learning rate must be finite");
    SD_CHECK(c.jac_valid, SDDDST_ERR_STATE, "no Jacobian assembled");
    SD_CHECK(c.spai, SDDDST_ERR_STATE, "no weights installed (sdddst_spai_train_init or set_weights)");
    if (!c.gm) c.gm = new Fgmres();
    int rc = c.gm->pair_factor(c);
    if (rc) return rc;
    return spai_train_step(c, *c.gm, probes, seed, eps, lr, loss, grad);
}

int sdddst_spai_graph(sdddst_ctx *ctx, int64_t *n_edges, int *Cp, int *Ci, double *Ce, double *Cd, float *gamma_e,
                      float *gamma_d, double *bd, double *bo, int *bcol)
{
    CTX_CHECK(ctx);
    SD_CHECK(n_edges, SDDDST_ERR_ARG, "n_edges is required");
    SD_CHECK(ctx->c.gm, SDDDST_ERR_STATE, "no learned preconditioner evaluated");
    return spai_graph(ctx->c, *ctx->c.gm, n_edges, Cp, Ci, Ce, Cd, gamma_e, gamma_d, bd, bo, bcol);
}

int sdddst_newton_update(sdddst_ctx *ctx, int slot, const double *delta_host)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(slot), SDDDST_ERR_ARG, "bad slot");
    if (delta_host) SD_HIP(hipMemcpyAsync(c.delta, delta_host, sizeof(double) * c.n, hipMemcpyHostToDevice, c.stream));
    return launch_newton_update(c, slot);
}

int sdddst_error_control(sdddst_ctx *ctx, double h, double hmax, int *accept, double *h_next, double *max_ratio,
                         int *selected_id)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    int rc = launch_error_control(c, h, hmax);
    if (rc) return rc;
    if ((rc = host_wait(c, c.ds, (int)(sizeof(DevScalars) / 8), c.hs))) return rc;
    if (accept) *accept = c.hs->accept;
    if (h_next) *h_next = c.hs->h_next;
    if (max_ratio) *max_ratio = c.hs->max_ratio;
    if (selected_id) *selected_id = c.hs->selected;
    return SDDDST_OK;
}

int sdddst_speed_sums(sdddst_ctx *ctx, int buf, double *sum_abs, double *sum_sq, double *sum_bv)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_buf(buf), SDDDST_ERR_ARG, "bad buffer");
    int rc = launch_speed_sums(c, buf);
    if (rc) return rc;
    if ((rc = host_wait(c, c.ds, (int)(sizeof(DevScalars) / 8), c.hs))) return rc;
    if (sum_abs) *sum_abs = c.hs->sum_abs;
    if (sum_sq) *sum_sq = c.hs->sum_sq;
    if (sum_bv) *sum_bv = c.hs->sum_bv;
    return SDDDST_OK;
}

int sdddst_strain_increment(sdddst_ctx *ctx, int old_slot, int new_slot, double *out)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(valid_slot(old_slot) && valid_slot(new_slot) && out, SDDDST_ERR_ARG, "bad arguments");
    int rc = launch_strain(c, old_slot, new_slot);
    if (rc) return rc;
    if ((rc = host_wait(c, c.ds, (int)(sizeof(DevScalars) / 8), c.hs))) return rc;
    *out = c.hs->strain;
    return SDDDST_OK;
}

int sdddst_stage_end(sdddst_ctx *ctx, double h, double hmax, int strain, double *vsq_init, double *vsq_init2,
                     int *accept, double *h_next, double *max_ratio, int *selected, double *strain_first,
                     double *strain_second)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    int rc;
    if ((rc = launch_speed_sums(c, SDDDST_V_INIT, c.ds + kStageSumA)) ||
        (rc = launch_speed_sums(c, SDDDST_V_INIT2, c.ds + kStageSumB)) || (rc = launch_error_control(c, h, hmax)))
        return rc;
    if (strain && ((rc = launch_strain(c, SDDDST_SLOT_DISLOCATIONS, SDDDST_SLOT_FIRST_SMALL, c.ds + kStageStrainA)) ||
                   (rc = launch_strain(c, SDDDST_SLOT_FIRST_SMALL, SDDDST_SLOT_SECOND_SMALL, c.ds + kStageStrainB))))
        return rc;
    if ((rc = launch_stage_pack(c))) return rc;
    double out[7];
    if ((rc = host_wait(c, c.ds + kStageOut, 7, out))) return rc;
    int ints[2];
    std::memcpy(ints, &out[4], sizeof ints);
    if (vsq_init) *vsq_init = out[0];
    if (vsq_init2) *vsq_init2 = out[1];
    if (max_ratio) *max_ratio = out[2];
    if (h_next) *h_next = out[3];
    if (accept) *accept = ints[0];
    if (selected) *selected = ints[1];
    if (strain_first) *strain_first = strain ? out[5] : 0.0;
    if (strain_second) *strain_second = strain ? out[6] : 0.0;
    return SDDDST_OK;
}

int sdddst_synchronize(sdddst_ctx *ctx)
{
    CTX_CHECK(ctx);
    SD_HIP(hipStreamSynchronize(ctx->c.stream));
    return SDDDST_OK;
}

int sdddst_field_eval(sdddst_ctx *ctx, int which, int64_t n, const double *dx, const double *dy, double *out)
{
    CTX_CHECK(ctx);
    return field_eval(ctx->c, which, n, dx, dy, out);
}

int sdddst_libm_eval(sdddst_ctx *ctx, int fn, int64_t n, const double *x, double *out)
{
    CTX_CHECK(ctx);
    return libm_eval(ctx->c, fn, n, x, out);
}

int sdddst_ev_decompose(sdddst_ctx *ctx, int cfg_slot, double *eigenvalues, double *eigenvectors, double *pn,
                        double *gan, double *sum_b)
{
    CTX_CHECK(ctx);
    SD_CHECK(valid_slot(cfg_slot), SDDDST_ERR_ARG, "bad slot");
    return ev_decompose(ctx->c, cfg_slot, eigenvalues, eigenvectors, pn, gan, sum_b);
}

int sdddst_ev_analysis(sdddst_ctx *ctx, int cfg_slot, const double *speeds, double stress_rate, double *eigenvalues,
                       double *eigenvectors, double *pn, double *gan, double *sum_b, double *deriv)
{
    CTX_CHECK(ctx);
    SD_CHECK(valid_slot(cfg_slot), SDDDST_ERR_ARG, "bad slot");
    SD_CHECK((speeds == nullptr) == (deriv == nullptr), SDDDST_ERR_ARG,
             "EV analysis: speeds and deriv go together");
    return ev_decompose(ctx->c, cfg_slot, eigenvalues, eigenvectors, pn, gan, sum_b, speeds, stress_rate, deriv);
}

int sdddst_timing(sdddst_ctx *ctx, int enable)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_HIP(hipStreamSynchronize(c.stream));
    SD_CHECK(enable >= 0 && enable <= 2, SDDDST_ERR_ARG, "timing mode %d (0 off, 1 all, 2 pair kernel)", enable);
    for (Context::Timer *t : {&c.t_pair, &c.t_speeds, &c.t_jac}) {
        t->used = 0;
        t->total_ms = 0;
        t->count = 0;
        t->calls = 0;
        t->every = 1;
    }
    if (enable == 2) c.t_pair.every = 4;
    c.timing = enable == 1 ? (kTimePair | kTimeSpeeds | kTimeJacobian) : enable == 2 ? kTimePair : 0;
    return SDDDST_OK;
}

int sdddst_kernel_stats(sdddst_ctx *ctx, const char *kernel, int *count, double *total_ms)
{
    CTX_CHECK(ctx);
    Context &c = ctx->c;
    SD_CHECK(kernel && count && total_ms, SDDDST_ERR_ARG, "bad arguments");
    Context::Timer *t = nullptr;
    if (strcmp(kernel, "pair") == 0) t = &c.t_pair;
    else if (strcmp(kernel, "speeds") == 0) t = &c.t_speeds;
    else if (strcmp(kernel, "jacobian") == 0) t = &c.t_jac;
    SD_CHECK(t, SDDDST_ERR_ARG, "unknown kernel name '%s'", kernel);
    int rc = timer_collect(c, *t);
    if (rc) return rc;
    *count = t->count;
    *total_ms = t->total_ms;
    return SDDDST_OK;
}

}  // extern "C"
