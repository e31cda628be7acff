// Device linear solvers for the Newton correction A delta = g.
#pragma once

#include <rocblas/rocblas.h>

#include "context.h"

namespace sdamd {

struct DenseLu {
    double *A = nullptr;
    int64_t *ipiv = nullptr;     // 64-bit interface: N^2 exceeds 2^31 above N = 46340
    int64_t *info = nullptr;
    rocblas_handle handle = nullptr;
    int init(Context &c);
    int factor(Context &c);
    int solve(Context &c);
    void release();
};

// Strong-row block (solver_strong.hip): a dense LU of A restricted to the
// rows H whose diagonal is far from 1 or that hold a large off-diagonal
// entry, applied on top of the pair-block inverse of the other rows.
struct StrongBlock {
    static constexpr int kMaxRows = 32768;   // 8 GiB of FP64 factor
    int n_alloc = 0, nh = 0, nh_alloc = 0;
    int *flag = nullptr;       // n: row in H
    int *pos = nullptr;        // n + 1: exclusive scan of flag (row index in H)
    long long *total = nullptr;
    double *hv = nullptr;      // n: right-hand side / solution on H
    double *D = nullptr;       // nh x nh column-major LU factors
    int64_t *piv = nullptr, *info = nullptr;
    rocblas_handle handle = nullptr;
    int factor(Context &c, double thr);
    int apply(Context &c, const double *v, double *z);   // z holds the pair-block result on entry
    void release_rows();
    void release();
};

// Restarted flexible GMRES (FP64) on the CSR view of A with a
// right preconditioner; see solver_fgmres.hip.
struct KrylovDev;    // state of a device-driven restart cycle (solver_fgmres.hip)
struct KrylovHost;   // what the host reads of it (mapped memory)

struct Fgmres {
    int n = 0, m = 0;
    double *V = nullptr;     // (m+1) x n Krylov basis
    double *Z = nullptr;     // m x n preconditioned directions
    double *w = nullptr;     // n
    double *dinv = nullptr;  // Jacobi: inverse diagonal of A
    double *red = nullptr;   // reduction scratch
    double *hcol = nullptr;  // device copy of one Hessenberg column
    double *H = nullptr;     // host copy of reduction results (mapped, coherent pinned memory)
    double *Hdev = nullptr;  // its device address
    unsigned *ticket = nullptr;  // last-block counter of the fused reductions
    double *binv = nullptr;      // pair-block inverse rows (2 per row)
    int *bcol = nullptr;
    KrylovDev *kst = nullptr;    // device cycle state
    KrylovHost *kh = nullptr;    // its host-read part (mapped, coherent pinned memory)
    KrylovHost *khdev = nullptr; // device address of kh
    double iter_mean = 3.0;      // running mean of iterations per solve (first batch size)
    StrongBlock strong;          // strong-row block of the current matrix
    bool strong_valid = false;   // factored for the current matrix
    int init(Context &c);
    int strong_factor(Context &c);                       // pair blocks + strong-row block, once per matrix
    int factor(Context &c);
    int pair_factor(Context &c);                         // pair-block inverse of A
    int apply(Context &c, const double *v, double *z);   // z = M^-1 v
    int solve(Context &c);
    void release();
};

// Learned sparse approximate inverse (spai.hip) on top of the pair-block
// inverse held by f.
int spai_param_count();
int spai_factor(Context &c, const Fgmres &f);
int spai_apply(Context &c, const Fgmres &f, const double *v, double *z);
int spai_set_weights(Context &c, const float *w);
int spai_get_weights(Context &c, float *w);
int spai_train_init(Context &c, unsigned long long seed);
int spai_train_step(Context &c, Fgmres &f, int k, unsigned long long seed, double eps, double lr, double *loss,
                    double *grad);
int spai_graph(Context &c, const Fgmres &f, int64_t *ne, int *Cp, int *Ci, double *Ce, double *Cd, float *ge,
               float *gd, double *bd, double *bo, int *bcol);

}  // namespace sdamd
