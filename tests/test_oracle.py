"""Pins of the CPU oracle (CPU only): against the reference's own translation
units (oracle/_ref, when built), against the committed golden vectors made
from them, and against what the mathematics fixes (image sums, symmetries,
limits, derivatives, exact identities of the Jacobian)."""
import ctypes
import os

import mpmath
import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


# -- golden vectors (made from the reference's own AnalyticField etc.) ----

def test_golden_field_points():
    g = np.load(os.path.join(GOLDEN, "field_points.npz"))
    xy, xyd = O.xy_batch(g["dx"], g["dy"])
    assert _bits_equal(xy, g["xy"])
    assert _bits_equal(xyd, g["xy_diff_x"])


def test_golden_field_points_dx2():
    g = np.load(os.path.join(GOLDEN, "field_points_dx2.npz"))
    lib = O.oracle()
    got = np.array([lib.orc_xy_diff_x2(float(u), float(v)) for u, v in zip(g["dx"], g["dy"])])
    assert _bits_equal(got, g["xy_diff_x2"])


def test_golden_speeds_reference_loop():
    g = np.load(os.path.join(GOLDEN, "speeds_n256_seed1.npz"))
    v0, _ = O.speeds(g["x"], g["y"], g["b"], sigma=0.0)
    v2, _ = O.speeds(g["x"], g["y"], g["b"], sigma=0.02)
    assert _bits_equal(v0, g["v_zero"])
    assert _bits_equal(v2, g["v_sigma002"])
    x, y, b = configs.dislocations(256, 1)
    assert _bits_equal(x, g["x"]) and _bits_equal(y, g["y"]) and _bits_equal(b, g["b"])


def test_golden_precision_handler():
    g = np.load(os.path.join(GOLDEN, "precision_handler.npz"))
    lib = O.oracle()
    ro = eo = 0
    for k in range(len(g["size"])):
        size, prec = int(g["size"][k]), float(g["prec"][k])
        r2 = g["r2"][ro:ro + g["r2_len"][k]]
        ids = g["ids"][ro:ro + g["r2_len"][k]]
        ro += g["r2_len"][k]
        err = g["err"][eo:eo + g["err_len"][k]]
        eo += g["err_len"][k]
        tol = np.full(size, prec * prec)
        for r, i in zip(r2, ids):
            lib.orc_ph_update_tolerance(O.dp(tol), prec * prec, float(r), int(i))
        sel = ctypes.c_int()
        mr = lib.orc_max_error_ratio(size, O.dp(err), O.dp(np.zeros(size)), O.dp(tol), ctypes.byref(sel))
        assert mr == g["max_ratio"][k]
        assert lib.orc_new_step_size(mr, float(g["old"][k])) == g["new_step"][k]


def test_golden_stress_protocols():
    g = np.load(os.path.join(GOLDEN, "stress_protocols.npz"))["table"]
    for kind, rate, spring, t, strain, typ, val in g:
        if kind == 0:
            expect = 0.0
        elif kind == 1:
            expect = t * rate
        else:
            expect = max(spring * (t * rate - strain), 0.0)
        assert expect == val


# -- against the reference build directly, when it is present ------------

@pytest.fixture(scope="module")
def ref():
    r = O.reference()
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return r


def test_reference_field_bit_identical(ref):
    rng = np.random.default_rng(3)
    n = 60000
    dx = np.concatenate([rng.uniform(-0.5, 0.5, n), rng.uniform(-2e-3, 2e-3, n // 3)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, n), rng.uniform(-2e-3, 2e-3, n // 3)])
    a, b = O.xy_batch(dx, dy)
    ra, rb = np.empty_like(dx), np.empty_like(dx)
    ref.ref_xy_batch(len(dx), O.dp(dx), O.dp(dy), O.dp(ra), O.dp(rb))
    assert _bits_equal(a, ra) and _bits_equal(b, rb)


def test_reference_xy_diff_x2_bit_identical(ref):
    rng = np.random.default_rng(5)
    lib = O.oracle()
    pts = np.concatenate([rng.uniform(-0.5, 0.5, (20000, 2)), rng.uniform(-2e-3, 2e-3, (4000, 2))])
    got = np.array([lib.orc_xy_diff_x2(u, v) for u, v in pts])
    want = np.array([ref.ref_xy_diff_x2(u, v) for u, v in pts])
    assert _bits_equal(got, want)


def test_reference_normalize(ref):
    lib = O.oracle()
    for v in np.concatenate([np.random.default_rng(4).uniform(-5, 5, 20000),
                             [0.5, -0.5, 1.5, -1.5, 0.49999999999999994, -0.5000000000000001, 2.0, -2.0, 0.0]]):
        assert lib.orc_normalize(v) == ref.ref_normalize(v)


def test_reference_speeds_loop(ref):
    x, y, b = configs.dislocations(300, 9)
    v, _ = O.speeds(x, y, b, sigma=0.01)
    vr = np.empty(300)
    ref.ref_speeds(300, O.dp(x), O.dp(y), O.dp(b), 0.01, O.dp(vr))
    assert _bits_equal(v, vr)


# -- mathematics ------------------------------------------------------------

mpmath.mp.dps = 40


def _wall_mp(x, y):
    """One wall of images along y, summed: sum_m x(x^2-(y+m)^2)/(x^2+(y+m)^2)^2."""
    x, y = mpmath.mpf(x), mpmath.mpf(y)
    return mpmath.nsum(lambda m: x * (x * x - (y + m) ** 2) / (x * x + (y + m) ** 2) ** 2, [-mpmath.inf, mpmath.inf])


def _xy_mp(dx, dy):
    """The reference's image structure with walls from the y-image sum."""
    f = lambda a: _wall_mp(a, dy)  # noqa: E731
    g = sum(f(dx + k) for k in range(-3, 4))
    d = mpmath.mpf(dx)
    if dx < 0:
        return -d * f(dx + 5) + (1 + d) * f(dx - 4) + f(dx + 4) + g
    return d * f(dx - 5) + (1 - d) * f(dx + 4) + f(dx - 4) + g


@pytest.mark.parametrize("dx,dy", [(0.1, 0.2), (-0.3, 0.05), (0.45, -0.4), (-0.02, 0.3), (0.2, 0.0),
                                   (0.003, -0.002), (-0.49, 0.49)])
def test_xy_equals_image_sum(dx, dy):
    v = O.oracle().orc_xy(dx, dy)
    ref = float(_xy_mp(dx, dy))
    scale = 1.0 + abs(ref)
    assert abs(v - ref) <= 1e-12 * scale


@pytest.mark.parametrize("dx,dy", [(0.1, 0.2), (-0.3, 0.05), (0.02, -0.31), (0.004, 0.003)])
def test_xy_diff_x_equals_derivative(dx, dy):
    d = float(mpmath.diff(lambda t: _xy_mp(t, dy), mpmath.mpf(dx)))
    v = O.oracle().orc_xy_diff_x(dx, dy)
    assert abs(v - d) <= 1e-9 * (1.0 + abs(d))


@pytest.mark.parametrize("dx,dy", [(0.1, 0.2), (-0.3, 0.05)])
def test_xy_diff_x2_against_wall_derivatives(dx, dy):
    # AnalyticField.cpp:95-106 leaves out the factor 2 pi^2 that f and f_dx
    # carry, so every f_dx2 term of xy_diff_x2 is the wall's second
    # derivative over 2 pi^2, while the two boundary terms 2 f_dx keep it:
    # with d2 the second derivative of the image structure and e its part
    # 2 (f'(dx-4) - f'(dx+5)) for dx < 0, 2 (f'(dx-5) - f'(dx+4)) otherwise,
    # the reference's value is (d2 - e) / (2 pi^2) + e; walls and their
    # derivatives from the y-image sum in mpmath
    F = lambda a: _wall_mp(a, dy)  # noqa: E731
    F1 = lambda a: mpmath.diff(F, mpmath.mpf(a), 1)  # noqa: E731
    d2 = mpmath.diff(lambda t: _xy_mp(t, dy), mpmath.mpf(dx), 2)
    extra = 2 * (F1(dx - 4) - F1(dx + 5)) if dx < 0 else 2 * (F1(dx - 5) - F1(dx + 4))
    want = float((d2 - extra) / (2 * mpmath.pi ** 2) + extra)
    v = O.oracle().orc_xy_diff_x2(dx, dy)
    assert abs(v - want) <= 1e-13 * (1.0 + abs(want))
    # the factor itself: 2 pi^2 xy_diff_x2 is the central difference of the
    # oracle's first derivative up to the boundary terms (1e-8 here)
    lib, h = O.oracle(), 1e-5
    fd = (lib.orc_xy_diff_x(dx + h, dy) - lib.orc_xy_diff_x(dx - h, dy)) / (2 * h)
    assert abs(v * 2.0 * np.pi ** 2 - fd) <= 1e-8 * (1.0 + abs(fd))


def test_xy_diff_x2_symmetries():
    # second x-derivative of a field odd in x and even in y: odd in x, even in y
    rng = np.random.default_rng(12)
    lib = O.oracle()
    for _ in range(2000):
        dx, dy = rng.uniform(-0.5, 0.5, 2)
        a = lib.orc_xy_diff_x2(dx, dy)
        assert abs(lib.orc_xy_diff_x2(-dx, dy) + a) <= 1e-11 * (1 + abs(a))
        assert abs(lib.orc_xy_diff_x2(dx, -dy) - a) <= 1e-11 * (1 + abs(a))


def test_xy_symmetries():
    rng = np.random.default_rng(11)
    lib = O.oracle()
    for _ in range(2000):
        dx, dy = rng.uniform(-0.5, 0.5, 2)
        a = lib.orc_xy(dx, dy)
        assert abs(lib.orc_xy(-dx, dy) + a) <= 1e-12 * (1 + abs(a))      # odd in x
        assert abs(lib.orc_xy(dx, -dy) - a) <= 1e-12 * (1 + abs(a))      # even in y


def test_xy_small_r_limit():
    lib = O.oracle()
    for r, ang in [(1e-4, 0.3), (5e-5, 1.2), (2e-4, 2.5)]:
        dx, dy = r * np.cos(ang), r * np.sin(ang)
        single = dx * (dx * dx - dy * dy) / (dx * dx + dy * dy) ** 2
        v = lib.orc_xy(dx, dy)
        assert abs(v - single) < 20.0      # image contributions are O(1); single is O(1/r)
        assert abs(single) > 1e3


def test_taylor_branch_continuity():
    lib = O.oracle()
    for ang in np.linspace(0.1, 3.0, 15):
        r_in, r_out = 0.99999999e-3, 1.00000001e-3
        vi = lib.orc_xy(r_in * np.cos(ang), r_in * np.sin(ang))
        vo = lib.orc_xy(r_out * np.cos(ang), r_out * np.sin(ang))
        assert abs(vi - vo) <= 1e-7 * abs(vo) + 1e-9


def test_dipole_and_sum_rule():
    lib = O.oracle()
    x, y, b = np.array([0.1, -0.2]), np.array([0.05, 0.3]), np.array([1.0, -1.0])
    v, _ = O.speeds(x, y, b, sigma=0.0)
    t = 1.0 * -1.0 * lib.orc_xy(lib.orc_normalize(0.1 - -0.2), lib.orc_normalize(0.05 - 0.3))
    assert v[0] == t and v[1] == -t
    x, y, b = configs.dislocations(400, 21)
    v, _ = O.speeds(x, y, b, sigma=0.03)
    assert abs(v.sum() - 0.03 * b.sum()) <= 1e-10 * np.abs(v).sum()


def test_speeds_against_long_double_sum():
    n = 200
    x, y, b = configs.dislocations(n, 5)
    lib = O.oracle()
    acc = np.zeros(n, dtype=np.longdouble)
    mag = np.zeros(n)
    for i in range(n):
        for j in range(i + 1, n):
            t = b[i] * b[j] * lib.orc_xy(lib.orc_normalize(x[i] - x[j]), lib.orc_normalize(y[i] - y[j]))
            acc[i] += t
            acc[j] -= t
            mag[i] += abs(t)
            mag[j] += abs(t)
    v, _ = O.speeds(x, y, b)
    assert np.all(np.abs(v - acc.astype(np.float64)) <= n * 2.3e-16 * mag)


def test_tolerance_rule():
    x, y, b = configs.dislocations(128, 2)
    _, tol = O.speeds(x, y, b, prec=1e-3)
    lib = O.oracle()
    for i in range(128):
        r2 = min((lib.orc_normalize(x[i] - x[j]) ** 2 + lib.orc_normalize(y[i] - y[j]) ** 2)
                 for j in range(128) if j != i)
        assert tol[i] == min(1e-6, r2 * 0.25 * 1e-2)


# -- Jacobian ----------------------------------------------------------------

def _predicate_numpy(x, y, rc, rc2):
    n = len(x)
    hi, lo = np.triu_indices(n, 1)[1], np.triu_indices(n, 1)[0]   # pairs lo < hi
    norm = np.vectorize(O.oracle().orc_normalize)
    dx = norm(x[hi] - x[lo])
    dy = norm(y[hi] - y[lo])
    s = np.sqrt(dx * dx + dy * dy) - rc
    keep = s * s < 36.8 * rc2
    return set(zip(lo[keep].tolist(), hi[keep].tolist()))


@pytest.mark.parametrize("n,m", [(256, 1.0), (200, 2.0), (60, 1.0)])
def test_jacobian_pattern_is_the_predicate(n, m):
    x, y, b = configs.dislocations(n, 13)
    Ap, Ai, Ax, dv, idx, raw = O.jacobian(x, y, b, m, 1e-4)
    rc, rc2, _ = O.cutoff(n, m)
    pairs = _predicate_numpy(x, y, rc, rc2)
    got = set()
    for j in range(n):
        rows = Ai[Ap[j]:Ap[j + 1]]
        assert np.all(np.diff(rows) > 0)                 # ascending rows
        assert rows[idx[j] - Ap[j]] == j                 # diagonal position
        for i in rows:
            if i != j:
                got.add((min(i, j), max(i, j)))
                # symmetric: (j, i) is in column i
                assert j in Ai[Ap[i]:Ap[i + 1]]
    assert got == pairs


def test_jacobian_value_identities():
    n, h = 256, 3e-4
    x, y, b = configs.dislocations(n, 17)
    Ap, Ai, Ax, dv, idx, raw = O.jacobian(x, y, b, 1.0, h)
    lib = O.oracle()
    for j in range(0, n, 17):
        col = slice(Ap[j], Ap[j + 1])
        rows, vals = Ai[col], raw[col]
        # pass 2: s = -(sequential column sum of the raw values), d from s
        s = 0.0
        for v in vals:
            s += v
        s *= -1.0
        d_expect = 1.0 / ((1.0 / s + 1.0) * (1.0 / s + 1.0)) if s > 0 else 0.0
        assert dv[j] == d_expect
        assert Ax[idx[j]] == s * ((1.0 + dv[j]) * 0.5) + 1.0
        for r, v in zip(rows, vals):
            if r > j:   # lower entries: h * b_r b_j * tau_x * taper
                dx, dy = lib.orc_normalize(x[r] - x[j]), lib.orc_normalize(y[r] - y[j])
                rr = np.sqrt(dx * dx + dy * dy)
                rc, rc2, inv = O.cutoff(n, 1.0)
                mult = 1.0 if dx * dx + dy * dy <= rc2 else np.exp(-(rr - rc) ** 2 * inv)
                assert abs(v - h * b[r] * b[j] * lib.orc_xy_diff_x(dx, dy) * mult) <= 1e-13 * abs(v)
    # final values: row scaling and +1 on the diagonal
    for j in range(0, n, 31):
        for k in range(Ap[j], Ap[j + 1]):
            if Ai[k] != j:
                assert Ax[k] == raw[k] * ((1.0 + dv[Ai[k]]) * 0.5)
    # raw values are symmetric: (i, j) and (j, i) hold the same number
    for j in range(0, n, 13):
        for k in range(Ap[j], Ap[j + 1]):
            i = Ai[k]
            m = Ap[i] + np.searchsorted(Ai[Ap[i]:Ap[i + 1]], j)
            if i != j:
                assert raw[m] == raw[k]


def test_dense_lu_solve_residual():
    n = 300
    x, y, b = configs.dislocations(n, 19)
    Ap, Ai, Ax, dv, idx, raw = O.jacobian(x, y, b, 1.0, 1e-3)
    rhs = np.random.default_rng(2).normal(size=n)
    sol = np.empty(n)
    assert O.oracle().orc_dense_solve_csc(n, O.ip(Ap), O.ip(Ai), O.dp(Ax), O.dp(rhs), O.dp(sol)) == 0
    A = np.zeros((n, n))
    for j in range(n):
        A[Ai[Ap[j]:Ap[j + 1]], j] = Ax[Ap[j]:Ap[j + 1]]
    assert np.linalg.norm(A @ sol - rhs) <= 1e-13 * np.linalg.norm(rhs)
    assert np.allclose(sol, np.linalg.solve(A, rhs), rtol=1e-12, atol=1e-14)


def test_cutoff_rule():
    # simulation_data.cpp:208-218: m >= sqrt(2N)/12 switches to an infinite cutoff
    rc, rc2, _ = O.cutoff(256, 1.0)
    assert rc == 1.0 / np.sqrt(256) * 1.0
    rc, rc2, _ = O.cutoff(72, 1.0)
    assert rc == 1.0 / np.sqrt(72) * 1e20


def test_step_size_rule():
    lib = O.oracle()
    assert lib.orc_new_step_size(0.0, 1e-4) == 2e-4
    assert lib.orc_new_step_size(1e-9, 1e-4) == 0.9 * 1e-4 * 2.0      # growth capped at 2
    m = 0.3
    assert lib.orc_new_step_size(m, 1e-4) == 0.9 * 1e-4 * (1.0 / np.sqrt(m)) ** (1.0 / 3.0)


def test_oracle_sim_fixed_step_bookkeeping():
    x, y, b = configs.dislocations(64, 23)
    s = O.OracleSim(x, y, b, cutoff_multiplier=1.0, prec=1e-6, h0=1e-6, stress="zero", calculate_strain=True)
    x0 = s.x()
    for _ in range(3):
        s.set_step_size(1e-6)
        assert s.step() == 1
    log = s.log()
    assert log.shape == (4, 12)
    assert abs(s.time - 3e-6) < 1e-18
    # strain column equals sum b (x_new - x_old) accumulated
    strain = float(np.sum(b * (s.x() - x0)))
    assert abs(log[-1, 9] - strain) <= 1e-12 * (abs(strain) + 1e-12)
    # energy is the trapezoid sum of v^2 (column 10) times h (column 11 grows)
    assert np.all(np.diff(log[:, 11]) > 0)
    s.close()


def test_config_generator():
    x, y, b = configs.dislocations(1000, 7)
    assert np.all((x >= -0.5) & (x < 0.5)) and np.all((y >= -0.5) & (y < 0.5))
    assert b.sum() == 0 and set(np.unique(b)) == {-1.0, 1.0}
    x2, y2, b2 = configs.dislocations(1000, 7)
    assert _bits_equal(x, x2) and _bits_equal(y, y2)
    sizes = configs.ensemble_sizes()
    assert len(sizes) == 256 and min(sizes) >= 64 and max(sizes) <= 2048
    # SplitMix64 known answer for seed 0 (first output)
    assert int(configs.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


# -- point defects (simulation.cpp:111-128, 269-325) ---------------------------

def _pd_speed_mp(dx, dy, A, K):
    X = lambda t: mpmath.sin(2 * mpmath.pi * t) * mpmath.mpf("0.5") / mpmath.pi  # noqa: E731
    X2 = lambda t: (1 - mpmath.cos(2 * mpmath.pi * t)) * mpmath.mpf("0.5") / mpmath.pi ** 2  # noqa: E731
    R = X2(dx) + X2(dy)
    E = mpmath.exp(-K * R)
    return -2 * A * X(dx) * X(dy) * ((1 - E) / R - K * E) / R


@pytest.mark.parametrize("dx,dy", [(0.003, 0.002), (-0.01, 0.004), (0.001, -0.0015), (0.02, 0.01)])
def test_point_defect_jacobian_term_is_force_derivative(dx, dy):
    # the Maple expression of the Jacobian diagonal is d/dx of the point-defect
    # speed term (per unit Burgers vector, no taper)
    n = 4096
    A, K = 1e-4 * 16.0 / np.sqrt(n), 1.65 * 1.65 * 1e6 / 256.0 * n
    d = float(mpmath.diff(lambda t: _pd_speed_mp(t, mpmath.mpf(dy), A, K), mpmath.mpf(dx)))
    v = O.oracle().orc_pd_diag_term(dx, dy, 1.0, A, K, 1.0)
    assert abs(v - d) <= 1e-9 * max(abs(d), 1e-12)


def test_point_defect_speed_term():
    # speeds with defects minus speeds without = the defect force, summed
    n, pc = 100, 40
    x, y, b = configs.dislocations(n, 41)
    px, py = configs.point_defects(n, pc, 41)
    A, K = 1e-4 * 16.0 / np.sqrt(n), 1.65 * 1.65 * 1e6 / 256.0 * n
    v1, _ = O.speeds(x, y, b, px=px, py=py, A=A, K=K)
    v0, _ = O.speeds(x, y, b)
    for i in range(0, n, 9):
        f = sum(float(_pd_speed_mp(mpmath.mpf(O.oracle().orc_normalize(x[i] - px[l])),
                                   mpmath.mpf(O.oracle().orc_normalize(y[i] - py[l])), A, K)) for l in range(pc))
        assert abs((v1[i] - v0[i]) - f * b[i]) <= 1e-9 * (abs(f) + 1e-6)
