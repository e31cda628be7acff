"""Mathematical pins of the oracle's stepping path (CPU only).

The reference's simulation.cpp needs umfpack.h and is not built here, so its
calculateG (simulation.cpp:133-193), integrate (62-84), Jacobian passes 2-3
(349-383) and three-stage step (435-650) are pinned by what the mathematics
of the scheme fixes, each check chosen so that a plausible slip fails it:

- the assembled matrix is the derivative of the residual with respect to the
  new positions (central differences at an infinite cutoff, where the
  reference's matrix is exact): catches row-for-column scaling, the sign of
  the diagonal sum, the +1, a dropped b_i b_j and the point-defect fold;
- the residual vanishes for a consistent step (weights of the old and new
  speeds sum to one);
- on a dipole's linear mode one stage is the theta-scheme with
  theta = (1+d)/2 on the new speeds: catches swapped weights;
- big step against two half steps shrinks 8x per halving (local error h^3):
  a dropped old-speed term gives 4x;
- a fixed-step trajectory converges to a DOP853 solution of the same ODE at
  second order;
- N = 1 closed forms (constant and linearly ramped stress): the stages move
  by exactly h*b*sigma, the error ratio is exactly 0 and the step doubles.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.optimize import brentq

import oracle_lib as O
from sdddst_amd import configs


def _dense(n, Ap, Ai, Ax):
    A = np.zeros((n, n))
    for j in range(n):
        A[Ai[Ap[j]:Ap[j + 1]], j] = Ax[Ap[j]:Ap[j + 1]]
    return A


@pytest.mark.parametrize("with_defects", [False, True])
def test_jacobian_is_residual_derivative(with_defects):
    # N = 48, cutoff multiplier 1 >= sqrt(2N)/12: infinite cutoff, no taper,
    # every pair in the pattern (simulation_data.cpp:211-214)
    n, pc = 48, 6
    x, y, b = configs.dislocations(n, 31)
    kw = {}
    if with_defects:
        px, py = configs.point_defects(n, pc, 31)
        kw = dict(px=px, py=py, A=1.6e-3 / np.sqrt(pc), K=1.65 ** 2 * 1e6 / 256.0 * n)
    h, sigma = 1e-2, 0.02
    sim = O.OracleSim(x, y, b, cutoff_multiplier=1.0, stress="constant", p1=sigma, **kw)
    v_old, _ = O.speeds(x, y, b, sigma=sigma, **kw)
    xn = x + 0.3 * h * v_old
    Ap, Ai, Ax, dv, idx, raw = O.jacobian(xn, y, b, 1.0, h, **kw)
    assert Ap[-1] == n * n
    A = _dense(n, Ap, Ai, Ax)
    # d spans [0, 0.64] at this h, so a row/column mix-up is visible
    assert dv.min() == 0.0 and dv.max() > 0.5
    eps = 1e-6
    fd = np.zeros((n, n))
    for j in range(n):
        g = []
        for k in (-2, -1, 1, 2):
            z = xn.copy()
            z[j] += k * eps
            g.append(sim.residual(h, z, x, dv, v_old))
        fd[:, j] = (g[0] - 8.0 * g[1] + 8.0 * g[2] - g[3]) / (12.0 * eps)
    sim.close()
    scale = np.abs(A - np.eye(n)).max()
    assert scale > 10.0
    assert np.abs(fd - A).max() <= 1e-7 * scale
    if with_defects:
        Ap0, Ai0, Ax0, _, idx0, _ = O.jacobian(xn, y, b, 1.0, h)
        # the defects move the diagonal by far more than the tolerance
        assert np.abs(Ax[idx] - Ax0[idx0]).max() > 1e-3 * scale


def test_residual_of_consistent_step_is_minus_h_v():
    # at x_new = x_old with v_new = v_old the weights (1+d)/2 and (1-d)/2 sum
    # to one, whatever d is
    n = 40
    x, y, b = configs.dislocations(n, 37)
    sigma, h = 0.02, 3e-3
    sim = O.OracleSim(x, y, b, cutoff_multiplier=1.0, stress="constant", p1=sigma)
    v, _ = O.speeds(x, y, b, sigma=sigma)
    dv = np.random.default_rng(1).uniform(0.0, 0.9, n)
    g = sim.residual(h, x, x, dv, v)
    sim.close()
    # rounding of x_new - a - x_old - b: a few ulps of |x| + h|v|
    assert np.all(np.abs(g + h * v) <= 4.5e-16 * (np.abs(x) + h * np.abs(v)))
    # while a sum of weights off by d would leave d h |v|
    assert np.max(dv * h * np.abs(v)) > 1e3 * np.max(np.abs(g + h * v))


def _dipole_equilibrium(dy):
    lib = O.oracle()
    d = brentq(lambda t: lib.orc_xy(t, dy), 0.03, 0.08, xtol=1e-17)
    return d, lib.orc_xy_diff_x(d, dy)


@pytest.mark.parametrize("s", [0.05, 0.3, 1.0, 3.0, 10.0, 100.0])
def test_dipole_mode_amplification_is_theta_scheme(s):
    # b = (+1, -1) one glide-plane offset 0.05 apart: d(dx)/dt = -2 tau(dx, dy)
    # has a stable zero near 45 degrees.  Linearised there, one stage maps
    # eps -> R eps with R = (1 - (1-theta) mu) / (1 + theta mu), mu = 2 h tau_x,
    # theta = (1 + d)/2 on the new speeds, (1 - theta) on the old ones.
    dy = 0.05
    dstar, tx = _dipole_equilibrium(dy)
    assert tx > 0
    h = s / tx
    eps0 = 1e-7
    x = np.array([dstar + eps0, 0.0])
    sim = O.OracleSim(x, np.array([dy, 0.0]), np.array([1.0, -1.0]), cutoff_multiplier=1.0, h0=h, stress="zero")
    xn, dv = sim.stage(h)
    sim.close()
    assert dv[0] == dv[1] and 0.0 < dv[0] < 1.0
    R = (xn[0] - xn[1] - dstar) / eps0
    theta, mu = (1.0 + dv[0]) / 2.0, 2.0 * h * tx
    Rp = (1.0 - (1.0 - theta) * mu) / (1.0 + theta * mu)
    assert abs(R - Rp) <= 2e-6
    # swapped weights would give (1 - theta mu)/(1 + (1-theta) mu)
    Rswap = (1.0 - theta * mu) / (1.0 + (1.0 - theta) * mu)
    if s >= 0.3:
        assert abs(R - Rswap) > 1e-2
    # A-stable, and L-stable in the limit: |R| < 1, R -> 0 for mu -> inf
    assert abs(R) < 1.0
    if s >= 100:
        assert abs(R) < 0.01
    # the glide of the pair's centre is conserved (zero stress)
    assert abs(xn[0] + xn[1] - x.sum()) <= 1e-16


def _separated_16(seed):
    rng = np.random.default_rng(seed)
    g = np.arange(4) / 4.0 - 0.375
    X, Y = np.meshgrid(g, g)
    x = X.ravel() + rng.uniform(-0.04, 0.04, 16)
    y = Y.ravel() + rng.uniform(-0.04, 0.04, 16)
    b = np.array([1.0, -1.0] * 8)
    rng.shuffle(b)
    return x, y, b


def test_local_error_order_three():
    # |x_big - x_small2| = C h^3 (1 + O(h)): 8x per halving of h
    x, y, b = _separated_16(5)
    errs = []
    for k in range(6):
        h = 4e-3 / 2 ** k
        s = O.OracleSim(x, y, b, cutoff_multiplier=1.0, prec=1.0, h0=h, stress="constant", p1=0.02)
        assert s.step() == 1
        errs.append(np.max(np.abs(s.big_x() - s.x())))
        s.close()
    ratio = np.array(errs[:-1]) / np.array(errs[1:])
    assert np.all(np.diff(ratio) > 0)             # approaching the asymptote from below
    assert 7.8 < ratio[-1] < 8.2


def test_fixed_step_trajectory_converges_second_order():
    x, y, b = _separated_16(5)
    sigma, T = 0.02, 2e-3

    def rhs(_, z):
        return O.speeds(z, y, b, sigma=sigma)[0]

    ref = solve_ivp(rhs, (0.0, T), x, method="DOP853", rtol=1e-13, atol=1e-15).y[:, -1]
    assert np.max(np.abs(ref - x)) > 1e-2          # the configuration moves
    errs = []
    for k in range(2, 7):
        m = 2 ** k
        s = O.OracleSim(x, y, b, cutoff_multiplier=1.0, prec=1.0, h0=T / m, stress="constant", p1=sigma)
        for _ in range(m):
            s.set_step_size(T / m)
            assert s.step() == 1
        assert abs(s.time - T) <= 1e-17
        errs.append(np.max(np.abs(s.x() - ref)))
        s.close()
    ratio = np.array(errs[:-1]) / np.array(errs[1:])
    assert np.all((ratio > 3.9) & (ratio < 4.1))


@pytest.mark.parametrize("bv", [1.0, -1.0])
def test_single_dislocation_constant_stress_closed_form(bv):
    h, sigma = 2.0 ** -20, 0.5
    s = O.OracleSim(np.array([0.125]), np.array([0.0]), np.array([bv]), cutoff_multiplier=1.0, prec=1e-6, h0=h,
                    stress="constant", p1=sigma)
    assert s.step() == 1
    assert s.x()[0] == 0.125 + h * sigma * bv
    assert s.big_x()[0] == 0.125 + h * sigma * bv
    assert s.time == h and s.failed_steps == 0
    log = s.log()
    assert log[-1, 3] == 0.0                     # max error ratio exactly zero
    assert s.step_size == 2.0 * h                # precision_handler.cpp:105-108
    assert log[-1, 4] == sigma                   # mean |v| = |b sigma|
    assert log[-1, 9] == h * sigma               # strain b * dx = h sigma
    s.close()


def test_single_dislocation_ramped_stress_closed_form():
    # sigma(t) = rate t: the trapezoid rule is exact for a linear integrand,
    # so stage I, II and III end where the exact solution does, which needs
    # the stage times t, t + h/2, t + h (simulation.cpp:144-186)
    h, rate = 2.0 ** -20, 2.0
    s = O.OracleSim(np.array([0.125]), np.array([0.0]), np.array([1.0]), cutoff_multiplier=1.0, prec=1e-6, h0=h,
                    stress="fixed-rate", p1=rate)
    for k in range(3):
        t0 = s.time
        x0 = s.x()[0]
        hk = s.step_size
        assert s.step() == 1
        exact = x0 + rate * (t0 * hk + 0.5 * hk * hk)
        assert s.x()[0] == exact and s.big_x()[0] == exact
        assert s.log()[-1, 3] == 0.0 and s.step_size == 2.0 * hk
    s.close()
