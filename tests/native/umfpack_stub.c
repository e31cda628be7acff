/*
 * A stand-in libumfpack for the tests (test infrastructure only).
 *
 * It exports the five UMFPACK entry points the simulator's plain solver path
 * calls (simulation.cpp:83, 388-395 of the reference), with UMFPACK's
 * documented signatures and the documented meaning of their arguments
 * (compressed-column A of n_row x n_col, 0-based Ap/Ai/Ax; sys UMFPACK_A = 0
 * solves A X = B; Control and Info may be NULL; opaque Symbolic / Numeric
 * objects freed by the matching free call, which sets the pointer to NULL).
 * The factorisation is a dense LU with partial pivoting on the host: the
 * answer of any direct solver to rounding, which is all the simulator
 * relies on.  It also records what it was given (the last matrix handed to
 * umfpack_di_numeric, call counts and the live object count) so that a test
 * can compare the hand-off with the oracle's matrix.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define UMFPACK_OK 0
#define UMFPACK_ERROR_out_of_memory (-1)
#define UMFPACK_ERROR_invalid_Numeric_object (-3)
#define UMFPACK_ERROR_invalid_Symbolic_object (-4)
#define UMFPACK_ERROR_argument_missing (-5)
#define UMFPACK_ERROR_n_nonpositive (-6)
#define UMFPACK_ERROR_invalid_system (-13)
#define UMFPACK_WARNING_singular_matrix 1
#define UMFPACK_A 0

typedef struct { int magic, n; } stub_symbolic;
typedef struct { int magic, n; double *lu; int *piv; } stub_numeric;

static int calls_symbolic, calls_numeric, calls_solve, live_objects;
static int last_n;
static long last_nnz;
static int *last_Ap, *last_Ai;
static double *last_Ax;

int umfpack_di_symbolic(int n_row, int n_col, const int Ap[], const int Ai[], const double Ax[], void **Symbolic,
                        const double Control[], double Info[])
{
    (void)Ax; (void)Control; (void)Info;
    calls_symbolic++;
    if (!Ap || !Ai || !Symbolic) return UMFPACK_ERROR_argument_missing;
    if (n_row <= 0 || n_col != n_row) return UMFPACK_ERROR_n_nonpositive;
    stub_symbolic *s = (stub_symbolic *)malloc(sizeof *s);
    if (!s) return UMFPACK_ERROR_out_of_memory;
    s->magic = 0x53594d;
    s->n = n_row;
    *Symbolic = s;
    live_objects++;
    return UMFPACK_OK;
}

static void record(int n, const int *Ap, const int *Ai, const double *Ax)
{
    long nnz = Ap[n];
    free(last_Ap); free(last_Ai); free(last_Ax);
    last_Ap = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    last_Ai = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    last_Ax = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    memcpy(last_Ap, Ap, sizeof(int) * (size_t)(n + 1));
    memcpy(last_Ai, Ai, sizeof(int) * (size_t)nnz);
    memcpy(last_Ax, Ax, sizeof(double) * (size_t)nnz);
    last_n = n;
    last_nnz = nnz;
}

int umfpack_di_numeric(const int Ap[], const int Ai[], const double Ax[], void *Symbolic, void **Numeric,
                       const double Control[], double Info[])
{
    (void)Control; (void)Info;
    calls_numeric++;
    stub_symbolic *s = (stub_symbolic *)Symbolic;
    if (!Ap || !Ai || !Ax || !Numeric) return UMFPACK_ERROR_argument_missing;
    if (!s || s->magic != 0x53594d) return UMFPACK_ERROR_invalid_Symbolic_object;
    const int n = s->n;
    record(n, Ap, Ai, Ax);
    stub_numeric *f = (stub_numeric *)calloc(1, sizeof *f);
    if (!f) return UMFPACK_ERROR_out_of_memory;
    f->magic = 0x4e554d;
    f->n = n;
    f->lu = (double *)calloc((size_t)n * (size_t)n, sizeof(double));
    f->piv = (int *)malloc(sizeof(int) * (size_t)n);
    if (!f->lu || !f->piv) {
        free(f->lu); free(f->piv); free(f);
        return UMFPACK_ERROR_out_of_memory;
    }
    double *a = f->lu;   /* row-major */
    for (int j = 0; j < n; ++j)
        for (int k = Ap[j]; k < Ap[j + 1]; ++k) a[(size_t)Ai[k] * n + j] += Ax[k];
    int status = UMFPACK_OK;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(a[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        f->piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)p * n + j];
                a[(size_t)p * n + j] = t;
            }
        const double d = a[(size_t)k * n + k];
        if (d == 0.0) { status = UMFPACK_WARNING_singular_matrix; continue; }
        for (int i = k + 1; i < n; ++i) {
            const double l = a[(size_t)i * n + k] / d;
            a[(size_t)i * n + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
        }
    }
    *Numeric = f;
    live_objects++;
    return status;
}

int umfpack_di_solve(int sys, const int Ap[], const int Ai[], const double Ax[], double X[], const double B[],
                     void *Numeric, const double Control[], double Info[])
{
    (void)Ap; (void)Ai; (void)Ax; (void)Control; (void)Info;
    calls_solve++;
    stub_numeric *f = (stub_numeric *)Numeric;
    if (sys != UMFPACK_A) return UMFPACK_ERROR_invalid_system;
    if (!X || !B) return UMFPACK_ERROR_argument_missing;
    if (!f || f->magic != 0x4e554d) return UMFPACK_ERROR_invalid_Numeric_object;
    const int n = f->n;
    const double *a = f->lu;
    for (int i = 0; i < n; ++i) X[i] = B[i];
    for (int k = 0; k < n; ++k)
        if (f->piv[k] != k) {
            double t = X[k];
            X[k] = X[f->piv[k]];
            X[f->piv[k]] = t;
        }
    for (int i = 0; i < n; ++i) {
        double s = X[i];
        for (int j = 0; j < i; ++j) s -= a[(size_t)i * n + j] * X[j];
        X[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = X[i];
        for (int j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * X[j];
        X[i] = s / a[(size_t)i * n + i];
    }
    return UMFPACK_OK;
}

void umfpack_di_free_symbolic(void **Symbolic)
{
    if (!Symbolic || !*Symbolic) return;
    free(*Symbolic);
    *Symbolic = NULL;
    live_objects--;
}

void umfpack_di_free_numeric(void **Numeric)
{
    if (!Numeric || !*Numeric) return;
    stub_numeric *f = (stub_numeric *)*Numeric;
    free(f->lu);
    free(f->piv);
    free(f);
    *Numeric = NULL;
    live_objects--;
}

/* Test access to the record. */
void stub_umfpack_counts(int *symbolic, int *numeric, int *solve, int *live)
{
    *symbolic = calls_symbolic;
    *numeric = calls_numeric;
    *solve = calls_solve;
    *live = live_objects;
}

long stub_umfpack_last_matrix(int *n, int *Ap, int *Ai, double *Ax, long cap)
{
    *n = last_n;
    if (!last_Ap) return -1;
    if (Ap) memcpy(Ap, last_Ap, sizeof(int) * (size_t)(last_n + 1));
    if (cap >= last_nnz) {
        if (Ai) memcpy(Ai, last_Ai, sizeof(int) * (size_t)last_nnz);
        if (Ax) memcpy(Ax, last_Ax, sizeof(double) * (size_t)last_nnz);
    }
    return last_nnz;
}
