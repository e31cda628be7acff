"""The device elementary functions and field (sdddst_amd/csrc/libm_compat.h,
field.h) built for the host (tests/native) against the host C library and
the oracle, bit for bit, on CPU.  The GPU tests repeat the comparison with
the device build."""
import numpy as np
import pytest

import oracle_lib as O

FN = {"cos": 0, "sin": 1, "exp": 2, "expm1": 3, "cosh": 4, "sinh": 5}


def _mc1(fn, x):
    mc = O.mathcheck()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    mc.mc_batch1(fn, len(x), O.dp(x), O.dp(out))
    return out


def _mc2(fn, dx, dy):
    mc = O.mathcheck()
    out = np.empty_like(dx)
    mc.mc_batch2(fn, len(dx), O.dp(dx), O.dp(dy), O.dp(out))
    return out


@pytest.mark.parametrize("name,lo,hi", [("cos", -3.1416, 3.1416), ("cos", -0.03, 0.03), ("exp", -40, 40),
                                        ("expm1", -0.35, 0.35), ("expm1", -35, 35), ("cosh", -35, 35),
                                        ("cosh", -0.35, 0.35), ("sinh", -35, 35), ("sin", -3.1416, 3.1416)])
def test_host_build_matches_glibc(name, lo, hi):
    x = np.random.default_rng(hash(name) % 1000).uniform(lo, hi, 300_000)
    got = _mc1(FN[name], x)
    want = O.libm_batch(FN[name], x)
    mism = np.count_nonzero(got.view(np.int64) != want.view(np.int64))
    assert mism <= (2 if name == "sin" else 0)


def test_thresholds_of_the_branches():
    # dense sampling around the branch points of cosh / sinh / expm1 / cos
    pts = []
    for t in (0.34657359027997264, 1.0397207708399179, 22.0, 0.855469, 2.426265, 0.126, 2 ** -27, 2 ** -28):
        pts.append(t + np.arange(-2000, 2000) * np.spacing(t))
    x = np.concatenate(pts + [-p for p in pts])
    for name in ("cosh", "sinh", "expm1", "exp", "cos"):
        xs = x[np.abs(x) < 4.0] if name == "cos" else x     # cos is used for |x| <= pi
        got = _mc1(FN[name], xs)
        x_ = xs
        want = O.libm_batch(FN[name], x_)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name


def test_normalize_equals_loop():
    x = np.concatenate([np.random.default_rng(1).uniform(-6, 6, 100_000),
                        [0.5, -0.5, 1.5, -1.5, 2.5, 0.49999999999999994, -0.5000000000000001, 0.0, -0.0, 3.0]])
    got = _mc1(6, x)
    want = O.libm_batch(8, x)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_field_exact_bit_identical_to_oracle():
    rng = np.random.default_rng(9)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 80_000), rng.uniform(-2e-3, 2e-3, 20_000)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 80_000), rng.uniform(-2e-3, 2e-3, 20_000)])
    got = _mc2(0, dx, dy)
    want, _ = O.xy_batch(dx, dy)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_field_fast_and_derivative():
    rng = np.random.default_rng(10)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 50_000), rng.uniform(-2e-3, 2e-3, 10_000)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 50_000), rng.uniform(-2e-3, 2e-3, 10_000)])
    xy, xyd = O.xy_batch(dx, dy)
    fast = _mc2(1, dx, dy)
    assert np.max(np.abs(fast - xy) / (np.abs(xy) + 1.0)) < 4e-15
    d = _mc2(2, dx, dy)
    assert np.mean(d.view(np.int64) == xyd.view(np.int64)) > 0.998
    assert np.max(np.abs(d - xyd) / np.maximum(np.abs(xyd), 1.0)) < 1e-13
    # production derivative: far images from one shared exp and batched
    # reciprocals; error relative to |tau_x| + 1, as for the fast field
    df = _mc2(3, dx, dy)
    assert np.max(np.abs(df - xyd) / (np.abs(xyd) + 1.0)) < 1e-14


def test_field_second_derivative():
    # the EV analysis' xy_diff_x2 (AnalyticField.cpp:183-202) against the
    # oracle: the same glibc-identical cosh, sinh and cos (cos(4 pi dy) over
    # the quadrants |p| <= 2 pi), integer powers through the correctly
    # rounded powi.  Where glibc's pow(x, 3.0) is an ulp off the correctly
    # rounded cube, the reference's numerator, which cancels as r -> 0
    # (no small-r branch), magnifies that ulp: ~1e-10 at r = 0.008, below
    # 1e-14 beyond r = 0.05
    rng = np.random.default_rng(13)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 20_000), rng.uniform(-2e-2, 2e-2, 4_000)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 20_000), rng.uniform(-2e-2, 2e-2, 4_000)])
    lib = O.oracle()
    want = np.array([lib.orc_xy_diff_x2(u, v) for u, v in zip(dx, dy)])
    got = _mc2(4, dx, dy)
    assert np.mean(got.view(np.int64) == want.view(np.int64)) > 0.995
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    assert rel[np.hypot(dx, dy) > 0.05].max() < 1e-14
    assert rel.max() < 1e-9


def test_cos_over_two_pi_matches_glibc():
    # cos(4 pi dy) of xy_diff_x2: |p| up to 2 pi, quadrant boundaries included
    rng = np.random.default_rng(14)
    p = np.concatenate([4 * np.pi * rng.uniform(-0.5, 0.5, 200_000),
                        np.array([np.pi, -np.pi, 2 * np.pi, -2 * np.pi, 1.25 * np.pi, 1.75 * np.pi, 2.426265,
                                  np.nextafter(1.25 * np.pi, 0), np.nextafter(1.75 * np.pi, 9)])])
    mc = O.mathcheck()
    got = np.array([mc.mc_cos_2pi(float(v)) for v in p])
    want = O.libm_batch(0, p)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_far_pair_term():
    # the pair kernel's far pairs (circular |dx| >= 0.1) take cos(2 pi dy) as
    # cos a_i cos a_j + sin a_i sin a_j and e^{+-2 pi dx} as products of
    # per-dislocation e^{+-2 pi x}: within 3.5e-14 of the reference's term
    # relative to |tau| + 1 (3e6-sample max observed 2.2e-14, 99.99th
    # percentile 9.2e-15, mean 2.8e-16; at |dx| >= 0.05 the max would be
    # 7.9e-14), and odd under swapping the two dislocations bit for bit, so a
    # pair's value does not depend on which one a tile holds as the target
    mc = O.mathcheck()
    rng = np.random.default_rng(12)
    n = 400_000
    xi, yi, xj, yj = (rng.uniform(-0.5, 0.5, n) for _ in range(4))
    xi[: n // 4] += rng.integers(-3, 4, n // 4)
    yi[: n // 4] += rng.integers(-3, 4, n // 4)
    out, swp = np.empty(n), np.empty(n)
    mc.mc_batch_pair_far(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(out))
    mc.mc_batch_pair_far(n, O.dp(xj), O.dp(yj), O.dp(xi), O.dp(yi), O.dp(swp))
    dx, dy = np.empty(n), np.empty(n)
    ex, ey = xi - xj, yi - yj
    mc.mc_batch1(6, n, O.dp(ex), O.dp(dx))
    mc.mc_batch1(6, n, O.dp(ey), O.dp(dy))
    ref, _ = O.xy_batch(dx, dy)
    far = np.abs(dx) >= 0.1
    assert far.mean() > 0.75
    err = np.abs(out - ref) / (np.abs(ref) + 1.0)
    assert err[far].max() < 3.5e-14
    assert err[far].mean() < 1e-15
    assert np.array_equal(out[far].view(np.int64), (-swp[far]).view(np.int64))


def test_far_pair_term_close_in_x_far_in_y():
    # pairs with |dx| < 0.1 (down to 1e-12) take the far form too when
    # cos(2 pi dy) <= 0.8: cosh(2 pi dx) - cos(2 pi dy) >= 0.2 bounds the
    # cancellation as for the x-far pairs (3.2e5-sample max observed 1.6e-14,
    # 99.99th percentile 1.2e-14), and the form stays odd bit for bit
    mc = O.mathcheck()
    rng = np.random.default_rng(13)
    n = 400_000
    xi, yi = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)
    ddx = np.concatenate([rng.uniform(-0.1, 0.1, n // 2),
                          rng.choice([-1.0, 1.0], n // 2) * 10.0 ** rng.uniform(-12, -1, n // 2)])
    ddy = rng.choice([-1.0, 1.0], n) * rng.uniform(0.1, 0.5, n)
    xj, yj = xi - ddx, yi - ddy
    xi[: n // 4] += rng.integers(-3, 4, n // 4)
    yj[: n // 4] += rng.integers(-3, 4, n // 4)
    out, swp = np.empty(n), np.empty(n)
    mc.mc_batch_pair_far(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(out))
    mc.mc_batch_pair_far(n, O.dp(xj), O.dp(yj), O.dp(xi), O.dp(yi), O.dp(swp))
    dx, dy = np.empty(n), np.empty(n)
    ex, ey = xi - xj, yi - yj
    mc.mc_batch1(6, n, O.dp(ex), O.dp(dx))
    mc.mc_batch1(6, n, O.dp(ey), O.dp(dy))
    ref, _ = O.xy_batch(dx, dy)
    sel = np.cos(2.0 * np.pi * dy) <= 0.8
    assert sel.mean() > 0.9 and np.all(np.abs(dx[sel]) < 0.1 + 1e-12)
    err = np.abs(out - ref) / (np.abs(ref) + 1.0)
    assert err[sel].max() < 3.5e-14
    assert err[sel].mean() < 1e-15
    assert np.array_equal(out[sel].view(np.int64), (-swp[sel]).view(np.int64))


def test_integer_powers_correctly_rounded():
    import mpmath
    mpmath.mp.prec = 200
    mc = O.mathcheck()
    rng = np.random.default_rng(2)
    for v in rng.uniform(-7e-3, 7e-3, 300):
        for n in (3, 4, 5, 6, 7):
            assert mc.mc_powi(float(v), n) == float(mpmath.mpf(float(v)) ** n)


def test_point_defect_term_host_build():
    mc = O.mathcheck()
    mc.mc_pd_force_dx.restype = __import__("ctypes").c_double
    mc.mc_pd_force_dx.argtypes = [__import__("ctypes").c_double] * 4
    rng = np.random.default_rng(6)
    n = 65536
    A, K = 1e-4 * 16.0 / np.sqrt(n), 1.65 * 1.65 * 1e6 / 256.0 * n
    worst = 0.0
    same = 0
    for _ in range(3000):
        dx, dy = rng.uniform(-0.02, 0.02, 2)
        a = mc.mc_pd_force_dx(dx, dy, A, K)
        r = O.oracle().orc_pd_diag_term(dx, dy, 1.0, A, K, 1.0)
        same += a == r
        worst = max(worst, abs(a - r) / max(abs(r), 1e-300))
    assert worst < 1e-14 and same > 0.97 * 3000
