"""Learned preconditioner on the device against the CPU restatement
(oracle/spai_gnn.py): core graph and pair-block inverse bit-exact, network
outputs, loss and weight gradient within FP32 tolerances, training lowers
the loss, and FGMRES with the learned preconditioner solves to 1e-10."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import spai_gnn as G  # noqa: E402

import sdddst_amd as S  # noqa: E402
from sdddst_amd import configs  # noqa: E402

pytestmark = pytest.mark.gpu


def _device_system(n, m, seed, h, learned_first=False):
    # learned_first: the learned preconditioner is selected before the
    # assembly, whose count pass then writes the within-cutoff bitmask the
    # core graph reads (else the core graph tests the build's positions)
    x, y, b = configs.dislocations(n, seed)
    c = S.Context(n, cutoff_multiplier=m)
    c.upload_state(0, x, y, b)
    if learned_first:
        c.set_preconditioner(S.PRECOND_LEARNED)
    Ap, Ai, Ax, dv, diag = c.jacobian_csc(0, h)
    rc, _ = c.cutoff()
    return c, G.system_from_jacobian(x, y, Ap, Ai, Ax, diag, rc)


@pytest.mark.parametrize("n,m,seed,h,learned_first", [(1024, 2.0, 31, 2e-5, False), (700, 4.0, 32, 5e-6, False),
                                                       (1024, 2.0, 34, 2e-5, True), (700, 4.0, 35, 5e-6, True)])
def test_device_network_and_gradient_match_oracle(n, m, seed, h, learned_first):
    c, s = _device_system(n, m, seed, h, learned_first)
    with c:
        rng = np.random.default_rng(seed)
        th = (G.init_params(7) + 0.05 * rng.standard_normal(G.NPARAM)).astype(np.float32)
        c.spai_train_init(7)
        assert np.array_equal(c.spai_get_weights(), G.init_params(7).astype(np.float32))
        c.spai_set_weights(th)
        loss, grad = c.spai_train_step(8, 1234, lr=0.0, want_grad=True)
        gr = c.spai_graph()
    theta = th.astype(np.float64)
    # pair-block inverse and core graph: bit-exact
    assert np.array_equal(gr["bcol"], s.bcol)
    assert np.array_equal(gr["bd"].view(np.int64), s.bd.view(np.int64))
    assert np.array_equal(gr["bo"].view(np.int64), s.bo.view(np.int64))
    assert np.array_equal(gr["Cp"], s.Cp) and np.array_equal(gr["Ci"], s.ci)
    assert (s.bcol >= 0).sum() > 0 and len(s.ci) > 0
    # network outputs (FP32 on the device, FP64 in the oracle)
    L, g, ge, gd = G.loss_and_grad(theta, s, 8, 1234)
    scale = max(np.abs(ge).max(), np.abs(gd).max())
    assert np.max(np.abs(gr["gamma_e"] - ge)) <= 2e-5 * scale
    assert np.max(np.abs(gr["gamma_d"] - gd)) <= 2e-5 * scale
    assert abs(loss - L) <= 1e-5 * abs(L)
    assert np.linalg.norm(grad - g) <= 1e-3 * np.linalg.norm(g)
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    assert np.max(np.abs(grad[big] - g[big]) / np.abs(g[big])) < 2e-2


def test_device_training_lowers_the_loss():
    c, s = _device_system(1024, 2.0, 33, 2e-5)
    with c:
        c.spai_train_init(5)
        l0, _ = c.spai_train_step(8, 1, lr=0.0)
        for t in range(40):
            c.spai_train_step(8, 100 + t, lr=3e-3)
        l1, _ = c.spai_train_step(8, 1, lr=0.0)
    assert l1 < 0.7 * l0, (l0, l1)


def test_learned_fgmres_matches_direct_solve():
    n, m, h = 1024, 2.0, 2e-5
    x, y, b = configs.dislocations(n, 34)
    rng = np.random.default_rng(3)
    th = (G.init_params(9) + 0.02 * rng.standard_normal(G.NPARAM)).astype(np.float32)
    out = {}
    for solver, precond in ((S.SOLVER_DENSE_LU, S.PRECOND_PAIR), (S.SOLVER_FGMRES, S.PRECOND_LEARNED),
                            (S.SOLVER_FGMRES, S.PRECOND_PAIR)):
        with S.Context(n, cutoff_multiplier=m) as c:
            c.upload_state(0, x, y, b)
            c.speeds(0, S.V_INIT, 0.0)
            c.speeds(0, S.V_INIT2, 0.0)
            c.jacobian_csc(0, h)
            c.spai_set_weights(th)
            c.solver_select(solver, 1e-10, 300, 40)
            c.set_preconditioner(precond)
            c.set_learned_gate(-1)     # the network itself, whatever the matrix
            g = c.residual(h, 0, 0, S.V_INIT, S.V_INIT2)
            c.solver_factor()
            d = c.solver_solve(download=True)
            it, rr, fb = c.solver_stats()
            out[(solver, precond)] = (d, it, rr, fb)
    d0 = out[(S.SOLVER_DENSE_LU, S.PRECOND_PAIR)][0]
    d1, it1, rr1, fb1 = out[(S.SOLVER_FGMRES, S.PRECOND_LEARNED)]
    assert fb1 == 0 and rr1 <= 1e-10 and it1 > 0
    assert np.max(np.abs(d1 - d0)) <= 1e-8 * np.max(np.abs(d0))
    with pytest.raises(S.SdddstError):
        with S.Context(64) as c:
            x2, y2, b2 = configs.dislocations(64, 1)
            c.upload_state(0, x2, y2, b2)
            c.jacobian_csc(0, 1e-6)
            c.solver_select(S.SOLVER_FGMRES, 1e-10, 100, 20)
            c.set_preconditioner(S.PRECOND_LEARNED)
            c.solver_factor()      # no weights installed


def _solve_iters(n, m, x, y, b, h, precond, weights):
    with S.Context(n, cutoff_multiplier=m) as c:
        c.upload_state(0, x, y, b)
        c.speeds(0, S.V_INIT, 0.02)
        c.speeds(0, S.V_INIT2, 0.02)
        c.jacobian_csc(0, h)
        c.spai_set_weights(weights)
        c.solver_select(S.SOLVER_FGMRES, 1e-10, 300, 40)
        c.set_preconditioner(precond)
        c.set_learned_gate(-1)         # the network itself, whatever the matrix
        c.residual(h, 0, 0, S.V_INIT, S.V_INIT2)
        c.solver_factor()
        c.solver_solve()
        it, rr, fb = c.solver_stats()
        return it, rr, fb, c.precond_fallbacks()


@pytest.mark.parametrize("m,seed", [(2.0, 7001), (4.0, 7002)])
def test_trained_network_gains_over_pair_blocks(m, seed):
    """The shipped network against the method without it (gamma = 0: the
    pair-block preconditioner) on the Newton systems a simulation at N=2048
    emits (seeds not used in training), at FGMRES relative residual 1e-10."""
    from sdddst_amd.simulation import Simulation
    from sdddst_amd.spai_train import DEFAULT_WEIGHTS, load_weights
    w = load_weights(DEFAULT_WEIGHTS)
    n = 2048
    x, y, b = configs.dislocations(n, seed)
    sim = Simulation(x, y, b, cutoff_multiplier=m, precision=1e-6, initial_step=1e-6, stress="constant",
                     stress_value=0.02, solver="fgmres", preconditioner="pair")
    snaps = []
    while sim.state().succesful_steps < 200:
        st = sim.state()
        if st.succesful_steps % 25 == 0 and (not snaps or snaps[-1][0] != st.succesful_steps):
            snaps.append((st.succesful_steps, sim.positions().copy(), st.step_size))
        sim.step()
    sim.close()
    pair, learned = [], []
    for _, xs, h in snaps:
        it_p, rr_p, fb_p, _ = _solve_iters(n, m, xs, y, b, h, S.PRECOND_PAIR, w)
        it_l, rr_l, fb_l, pf_l = _solve_iters(n, m, xs, y, b, h, S.PRECOND_LEARNED, w)
        assert fb_p == 0 and fb_l == 0 and pf_l == 0 and rr_l <= 1e-10
        pair.append(it_p)
        learned.append(it_l)
    print(f"m={m}: FGMRES iterations pair {pair} learned {learned}")
    assert sum(learned) <= 0.85 * sum(pair)
    assert all(l <= p + 1 for l, p in zip(learned, pair))


def test_learned_gate_follows_the_shifted_diagonals():
    """With the learned preconditioner selected, a Newton matrix with more
    than max(16, N/256) columns whose scaled diagonal is more than 0.25 from
    1 (a step size large for the configuration) is preconditioned by the pair
    blocks alone, one with fewer by the network; the count is taken on the
    downloaded matrix here and the choice read from the solver history (1:
    pair blocks, 2: learned).  A stage-II style reuse at h / 2 scales the
    count by 1/2; gate -1 keeps the network at any step size.  Every solve
    still reaches 1e-10."""
    from sdddst_amd.spai_train import DEFAULT_WEIGHTS, load_weights
    w = load_weights(DEFAULT_WEIGHTS)
    n, m = 4096, 2.0
    x, y, b = configs.dislocations(n, 9)
    gate = max(16, n // 256)
    seen = set()
    with S.Context(n, cutoff_multiplier=m) as c:
        c.upload_state(0, x, y, b)
        c.speeds(0, S.V_INIT, 0.02)
        c.speeds(0, S.V_INIT2, 0.02)
        c.spai_set_weights(w)
        c.solver_select(S.SOLVER_FGMRES, 1e-10, 300, 40)
        c.set_preconditioner(S.PRECOND_LEARNED)
        c.set_direct_limit(0)
        for h in (1e-6, 1e-7, 1e-8, 1e-9):
            Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h)
            count = int(np.count_nonzero(np.abs(Ax[dg] - 1.0) > 0.25))
            for gate_set, want in ((-2, 2 if count <= gate else 1), (-1, 2)):
                c.set_learned_gate(gate_set)
                c.residual(h, 0, 0, S.V_INIT, S.V_INIT2)
                c.solver_factor()
                c.solver_history(1024)
                c.solver_solve()
                hist, _ = c.solver_history(1024)
                assert int(hist[0, 0]) == want, (h, count, gate_set, hist)
                assert hist[-1, 4] == 0 and hist[-1, 2] <= 1e-10
                if gate_set == -2:
                    seen.add(want)
        # the reuse at h / 2 of the last full build (rebuild = 2, stage II):
        # the count of that build, halved
        c.set_learned_gate(-2)
        for h_full in (1e-6, 1e-8):
            Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h_full)
            full = int(np.count_nonzero(np.abs(Ax[dg] - 1.0) > 0.25))
            nnz = ctypes.c_int64()
            c._check(c.lib.sdddst_jacobian_csc(c.h, 0, 0.5 * h_full, 2, None, None, None, None, None,
                                               ctypes.byref(nnz)))
            c.residual(0.5 * h_full, 0, 0, S.V_INIT, S.V_INIT2)
            c.solver_factor()
            c.solver_history(1024)
            c.solver_solve()
            hist, _ = c.solver_history(1024)
            assert int(hist[0, 0]) == (2 if 0.5 * full <= gate else 1), (h_full, full, hist)
    assert seen == {1, 2}, "the step sizes should cover both choices"
