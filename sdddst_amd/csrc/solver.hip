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
        if (rc == SDDDST_ERR_SOLVER && c.precond == SDDDST_PRECOND_LEARNED) {
            // the pair blocks are part of the learned preconditioner's factor
            c.precond_fallbacks++;
            c.precond_override = SDDDST_PRECOND_PAIR;
            rc = c.gm->solve(c);
            c.precond_override = -1;
            c.total_iters += c.last_iters;
        }
        if (rc == SDDDST_ERR_SOLVER) {
            // no convergence: fall back to the direct solver
            c.fallbacks++;
            if (!c.lu) c.lu = new DenseLu();
            rc = c.lu->factor(c);
            if (rc) return rc;
            return c.lu->solve(c);
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
