// Solver dispatch: the factor / solve pair the Newton loop calls once per
// integrate and once per iteration (simulation.cpp:66, 77).
#include "context.h"
#include "solvers.h"

namespace sdamd {

int solver_factor(Context &c)
{
    SD_CHECK(c.jac_valid, SDDDST_ERR_STATE, "no Jacobian assembled");
    if (c.solver == SDDDST_SOLVER_DENSE_LU) {
        if (!c.lu) c.lu = new DenseLu();
        return c.lu->factor(c);
    }
    if (c.solver == SDDDST_SOLVER_FGMRES) {
        // The Krylov solver works on the CSR view of A, which needs the
        // symmetric pattern; after a zero-value compaction the matrix goes
        // to the direct solver.
        c.direct_once = !c.jac_symmetric;
        if (c.direct_once) {
            if (!c.lu) c.lu = new DenseLu();
            return c.lu->factor(c);
        }
        if (!c.gm) c.gm = new Fgmres();
        return c.gm->factor(c);
    }
    set_error("solver %d has no device factorisation", c.solver);
    return SDDDST_ERR_UNSUPPORTED;
}

int solver_solve(Context &c)
{
    if (c.solver == SDDDST_SOLVER_DENSE_LU) {
        SD_CHECK(c.lu, SDDDST_ERR_STATE, "solve before factor");
        return c.lu->solve(c);
    }
    if (c.solver == SDDDST_SOLVER_FGMRES && c.direct_once) return c.lu->solve(c);
    if (c.solver == SDDDST_SOLVER_FGMRES) {
        SD_CHECK(c.gm, SDDDST_ERR_STATE, "solve before factor");
        int rc = c.gm->solve(c);
        c.total_iters += c.last_iters;
        c.total_solves += 1;
        c.log_solve(c.precond_active, c.last_iters, c.last_rel_res, c.last_gnorm, rc == SDDDST_OK ? 0 : 1);
        const int pair = c.precond_active == SDDDST_PRECOND_JACOBI ? SDDDST_PRECOND_JACOBI : SDDDST_PRECOND_PAIR;
        if (rc == SDDDST_ERR_SOLVER && c.precond_active == SDDDST_PRECOND_LEARNED) {
            // the pair blocks are part of the learned preconditioner's factor
            c.precond_fallbacks++;
            c.precond_override = pair;
            rc = c.gm->solve(c);
            c.precond_override = -1;
            c.total_iters += c.last_iters;
            c.log_solve(pair, c.last_iters, c.last_rel_res, c.last_gnorm, rc == SDDDST_OK ? 0 : 1);
        }
        if (rc == SDDDST_ERR_SOLVER && c.precond_active != SDDDST_PRECOND_STRONG) {
            // the pair blocks stagnate on indefinite systems (a step size
            // far too large for the configuration): once more with the
            // strong-row block, factored once for this matrix, before the
            // direct solver
            rc = c.gm->strong_factor(c);
            if (rc == SDDDST_OK) {
                c.precond_override = SDDDST_PRECOND_STRONG;
                rc = c.gm->solve(c);
                c.precond_override = -1;
                c.total_iters += c.last_iters;
                c.log_solve(SDDDST_PRECOND_STRONG, c.last_iters, c.last_rel_res, c.last_gnorm,
                            rc == SDDDST_OK ? 0 : 1, c.gm->strong.nh);
            } else {
                // too many strong rows (or the block is singular): the
                // direct solver decides
                c.log_solve(SDDDST_PRECOND_STRONG, 0, 1.0, c.last_gnorm, 1, c.gm->strong.nh);
                rc = SDDDST_ERR_SOLVER;
            }
            c.strong_solves++;
        }
        if (rc == SDDDST_ERR_SOLVER) {
            // no convergence: fall back to the direct solver, which holds the
            // matrix dense (8 N^2 bytes) and so is allowed only up to a size
            SD_CHECK(c.n <= c.direct_max_n, SDDDST_ERR_NOT_CONVERGED,
                     "FGMRES did not reach the relative residual %.1e (%.3e after %d iterations) and N = %d is "
                     "above the dense direct solver's limit %d", c.solver_tol, c.last_rel_res, c.last_iters, c.n,
                     c.direct_max_n);
            c.fallbacks++;
            if (!c.lu) c.lu = new DenseLu();
            rc = c.lu->factor(c);
            if (rc) return rc;
            rc = c.lu->solve(c);
            c.log_solve(-1, 0, 0.0, c.last_gnorm, 2);
            return rc;
        }
        return rc;
    }
    set_error("solver %d has no device solve", c.solver);
    return SDDDST_ERR_UNSUPPORTED;
}

void solver_free(Context &c)
{
    if (c.lu) {
        c.lu->release();
        delete c.lu;
        c.lu = nullptr;
    }
    if (c.gm) {
        c.gm->release();
        delete c.gm;
        c.gm = nullptr;
    }
}

}  // namespace sdamd
