"""The large combined workload (N = 65536 dislocations + 2048 point
defects, ramped stress, m = 2, learned-preconditioned FGMRES) and its solver
path: the point-defect terms with the learned FGMRES against the oracle at a
size the oracle steps in seconds; the strong-row block preconditioner on the
indefinite systems a far too large step produces; and, at full size,
properties that do not need the oracle's O(N^2) loops (determinism, the
Jacobian pattern on sampled columns against the predicate of
simulation.cpp:335, symmetry, and the first attempts from h = 1e-6 without a
dense N x N factorisation)."""
import numpy as np
import pytest

import oracle_lib as O
import sdddst_amd as S
from sdddst_amd import configs
from sdddst_amd._capi import solver_history
from sdddst_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def _pd_constants(n):
    # project_parser.cpp:208-212: A *= 1/sqrt(N), KASQR *= N
    return (1e-4 * 16.0) * (1.0 / np.sqrt(n)), 1.65 * 1.65 * 1e6 / 256.0 * n


def test_learned_fgmres_with_point_defects_matches_oracle():
    # config-5 physics at N = 512 + 96 defects, fixed-rate stress, learned
    # preconditioner, adaptive steps from h = 1e-6
    n, pc = 512, 96
    x, y, b = configs.dislocations(n, 5)
    px, py = configs.point_defects(n, pc, 5)
    A, K = _pd_constants(n)
    o = O.OracleSim(x, y, b, cutoff_multiplier=2.0, prec=1e-6, h0=1e-6, stress="fixed-rate", p1=200.0,
                    calculate_strain=True, px=px, py=py, A=A, K=K)
    g = Simulation(x, y, b, cutoff_multiplier=2.0, precision=1e-6, initial_step=1e-6, stress="fixed-rate",
                   stress_value=200.0, calculate_strain=True, solver="fgmres", preconditioner="learned",
                   field_mode="exact", px=px, py=py, A=A, KASQR=K, solver_tol=1e-10, solver_max_iter=300,
                   solver_restart=40)
    for _ in range(12):
        assert o.step() == g.step()
    st = g.state()
    assert st.succesful_steps == o.succesful_steps and st.failed_steps == o.failed_steps
    assert st.succesful_steps >= 5
    assert abs(st.sim_time - o.time) <= 1e-9 * o.time
    assert np.max(np.abs(g.positions() - o.x())) < 1e-9
    assert st.solver_fallbacks == 0
    o.close()
    g.close()


def _indefinite_system(n, pc, h, seed=5):
    # a step size ~100x what the configuration allows: A = I + (1+d)/2 M is
    # indefinite with a condition number ~1e6 (pinned dislocations, close
    # pairs); the residual g comes from a random perturbation of the positions
    x, y, b = configs.dislocations(n, seed)
    px, py = configs.point_defects(n, pc, seed)
    A, K = _pd_constants(n)
    c = S.Context(n, pc=pc, cutoff_multiplier=2.0, A=A, KASQR=K)
    c.upload_state(0, x, y, b)
    c.upload_points(px, py)
    c.jacobian_nnz(0, h)
    c.upload_speed(S.V_INIT, np.zeros(n))
    c.upload_state(S.SLOT_SCRATCH_A, x + np.random.default_rng(seed).normal(0, 1e-6, n))
    c.upload_state(S.SLOT_SCRATCH_B, x)
    gd = c.residual(h, S.SLOT_SCRATCH_A, S.SLOT_SCRATCH_B, S.V_INIT, S.V_INIT, download=True)
    return c, gd


def test_strong_block_solves_indefinite_system():
    n = 4096
    c, gd = _indefinite_system(n, 128, 1.6e-5)
    c.solver_select(S.SOLVER_DENSE_LU)
    c.solver_factor()
    d_lu = c.solver_solve(download=True)
    c.solver_select(S.SOLVER_FGMRES, 1e-10, 300, 40)
    c.set_preconditioner(S.PRECOND_STRONG)
    c.solver_factor()
    d = c.solver_solve(download=True)
    it, rr, fb = c.solver_stats()
    print(f"strong block: {it} iterations, relative residual {rr:.2e}")
    assert fb == 0 and rr <= 1e-10 and it < 150
    assert np.linalg.norm(d - d_lu) <= 1e-6 * np.linalg.norm(d_lu)
    # the pair blocks alone stagnate on this system; their solve is retried
    # with the strong-row block instead of the direct solver
    c.set_preconditioner(S.PRECOND_PAIR)
    c.solver_factor()
    solver_history(c.h)
    d2 = c.solver_solve(download=True)
    rec, strong = solver_history(c.h)
    it, rr, fb = c.solver_stats()
    assert fb == 0 and rr <= 1e-10 and strong >= 1
    assert rec[-1][0] == S.PRECOND_STRONG and rec[-1][4] == 0          # converged with the strong block
    assert np.linalg.norm(d2 - d_lu) <= 1e-6 * np.linalg.norm(d_lu)
    c.close()


def test_direct_fallback_size_limit_is_an_error():
    # with the strong-row block unable to help (limit forced to 0 rows is
    # not settable; a one-iteration budget at an unreachable tolerance is) and
    # N above the direct solver's limit, the solve reports an error instead of
    # factoring N x N
    n = 1024
    x, y, b = configs.dislocations(n, 8)
    with S.Context(n, cutoff_multiplier=1.0) as c:
        c.upload_state(0, x, y, b)
        c.jacobian_nnz(0, 5e-3)
        c.upload_speed(S.V_INIT, np.zeros(n))
        c.upload_state(S.SLOT_SCRATCH_A, x + 1e-3)
        c.upload_state(S.SLOT_SCRATCH_B, x)
        c.residual(1e-4, S.SLOT_SCRATCH_A, S.SLOT_SCRATCH_B, S.V_INIT, S.V_INIT)
        c.solver_select(S.SOLVER_FGMRES, 1e-14, 1, 1)
        c.set_direct_limit(512)
        c.solver_factor()
        with pytest.raises(S.SdddstError, match="limit"):
            c.solver_solve()
        c.set_direct_limit(1 << 20)
        c.solver_factor()
        c.solver_solve()
        assert c.solver_stats()[2] == 1


def _predicate_columns(x, y, cols, rc, rc2):
    out = {}
    for j in cols:
        dx = x - x[j]
        dy = y - y[j]
        # normalize by the loop semantics (|d| < 1 here: one step at most)
        dx = np.where(dx < -0.5, dx + 1.0, np.where(dx >= 0.5, dx - 1.0, dx))
        dy = np.where(dy < -0.5, dy + 1.0, np.where(dy >= 0.5, dy - 1.0, dy))
        s = np.sqrt(dx * dx + dy * dy) - rc
        keep = s * s < 36.8 * rc2
        keep[j] = True
        out[j] = np.nonzero(keep)[0]
    return out


def test_full_size_combined_workload():
    n, pc = 65536, 2048
    x, y, b = configs.dislocations(n, 5)
    px, py = configs.point_defects(n, pc, 5)
    A, K = _pd_constants(n)
    with S.Context(n, pc=pc, cutoff_multiplier=2.0, A=A, KASQR=K) as c:
        c.upload_state(0, x, y, b)
        c.upload_points(px, py)
        v1 = c.speeds(0, S.V_INIT, 0.05, update_tol=False)
        v2 = c.speeds(0, S.V_INIT2, 0.05, update_tol=False)
        assert np.array_equal(v1.view(np.int64), v2.view(np.int64))
        assert np.all(np.isfinite(v1))
        Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, 1e-6)
    nnz = int(Ap[-1])
    assert nnz == len(Ai) and np.all(np.diff(Ap) > 0)
    # symmetric pattern: (i, j) stored iff (j, i) stored
    cols = np.repeat(np.arange(n), np.diff(Ap))
    key = np.sort(cols.astype(np.int64) * n + Ai)
    keyT = np.sort(Ai.astype(np.int64) * n + cols)
    assert np.array_equal(key, keyT)
    # 64 sampled columns against the predicate, rows ascending, diagonal found
    rc, rc2, _ = O.cutoff(n, 2.0)
    sample = np.random.default_rng(1).choice(n, 64, replace=False)
    want = _predicate_columns(x, y, sample, rc, rc2)
    for j in sample:
        rows = Ai[Ap[j]:Ap[j + 1]]
        assert np.array_equal(rows, want[j])
        assert rows[dg[j] - Ap[j]] == j
    # the first attempts from h = 1e-6 (four orders of magnitude above what
    # this configuration allows): the Newton systems are indefinite; an
    # attempt whose solve does not converge is rejected with h / 8, never
    # through an N x N factorisation, until the steps succeed
    g = Simulation(x, y, b, cutoff_multiplier=2.0, precision=1e-6, initial_step=1e-6, stress="fixed-rate",
                   stress_value=4.0e6, solver="fgmres", preconditioner="learned", px=px, py=py, A=A, KASQR=K,
                   solver_tol=1e-10, solver_max_iter=300, solver_restart=40)
    hs = []
    for _ in range(14):
        hs.append(g.state().step_size)
        g.step()
        if g.state().succesful_steps >= 3:
            break
    st = g.state()
    rec, strong = solver_history(g.device_context)
    print(f"{len(hs)} attempts: {st.succesful_steps} successful, aborted {st.aborted_attempts}, "
          f"strong-row retries {strong}, step sizes {hs}")
    assert st.succesful_steps == 3
    assert st.solver_fallbacks == 0
    assert np.all(rec[:, 0] >= 0)                    # no direct solve
    for k in range(st.aborted_attempts):             # each aborted attempt divides h by 8
        assert hs[k + 1] == hs[k] * 0.125
    assert st.failed_steps >= st.aborted_attempts
    assert np.all(np.isfinite(g.positions()))
    g.close()
