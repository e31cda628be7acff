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


def test_cos_first_range_matches_glibc():
    # cos_r1, glibc's first cos range alone (the near pairs' |2 pi dy| <
    # 0.6436), bit for bit with glibc's cos below 0.855469, at random
    # arguments, near the range's end and below 2^-27
    rng = np.random.default_rng(21)
    t = 0.855469
    x = np.concatenate([rng.uniform(-t, t, 400_000), rng.uniform(-0.05, 0.05, 100_000),
                        np.nextafter(t, 0) - np.arange(4000) * np.spacing(t),
                        2.0 ** -27 + np.arange(-2000, 2000) * np.spacing(2.0 ** -27), [0.0, -0.0]])
    x = np.concatenate([x, -x])
    x = x[np.abs(x) < float.fromhex("0x1.b6p-1")]    # hiword below 0x3feb6000
    got = _mc1(11, x)
    want = O.libm_batch(0, x)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    # outside the range the lookup stays in the table (clamped), finite
    assert np.all(np.isfinite(_mc1(11, np.linspace(-4.0, 4.0, 10_001))))


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


def test_far_pair_term_by_gap():
    # the pair kernel takes the far form for every pair whose gap
    # cosh(2 pi dx) - cos(2 pi dy), as the far form computes it, is >= 0.2,
    # whatever |dx| (down to 1e-12): the gap bounds the cancellation as for
    # the x-far pairs (2.4e5-sample max observed 1.6e-14).  The gap is
    # symmetric in the two dislocations bit for bit, the term odd, and every
    # far pair has r^2 > 0.0081 (the far form is enabled only when no
    # tolerance candidate can be that far)
    mc = O.mathcheck()
    rng = np.random.default_rng(13)
    n = 400_000
    xi, yi = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)
    r = rng.uniform(0.06, 0.3, n)
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    ddx, ddy = r * np.cos(th), r * np.sin(th)
    ddx[: n // 8] = rng.choice([-1.0, 1.0], n // 8) * 10.0 ** rng.uniform(-12, -2, n // 8)
    xj, yj = xi - ddx, yi - ddy
    xi[: n // 4] += rng.integers(-3, 4, n // 4)
    yj[n // 4: n // 2] += rng.integers(-3, 4, n // 4)
    out, swp, gap, gsw = np.empty(n), np.empty(n), np.empty(n), np.empty(n)
    mc.mc_batch_pair_far(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(out))
    mc.mc_batch_pair_far(n, O.dp(xj), O.dp(yj), O.dp(xi), O.dp(yi), O.dp(swp))
    mc.mc_batch_pair_gap(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(gap))
    mc.mc_batch_pair_gap(n, O.dp(xj), O.dp(yj), O.dp(xi), O.dp(yi), O.dp(gsw))
    assert np.array_equal(gap.view(np.int64), gsw.view(np.int64))
    dx, dy = np.empty(n), np.empty(n)
    ex, ey = xi - xj, yi - yj
    mc.mc_batch1(6, n, O.dp(ex), O.dp(dx))
    mc.mc_batch1(6, n, O.dp(ey), O.dp(dy))
    exact_gap = 2.0 * np.sinh(np.pi * dx) ** 2 + 2.0 * np.sin(np.pi * dy) ** 2
    assert np.max(np.abs(gap - exact_gap)) < 1e-13
    sel = gap >= 0.2
    assert 0.5 < sel.mean() < 0.95 and (sel & (np.abs(dx) < 0.1)).sum() > 50_000
    assert np.all(dx[sel] ** 2 + dy[sel] ** 2 > 0.0081)
    ref, _ = O.xy_batch(dx, dy)
    err = np.abs(out - ref) / (np.abs(ref) + 1.0)
    assert err[sel].max() < 3.5e-14
    assert err[sel].mean() < 1e-15
    assert np.array_equal(out[sel].view(np.int64), (-swp[sel]).view(np.int64))


def test_mixed_pair_term():
    # the pairs of the pair kernel's mixed tiles: a far pair (gap >= 0.2)
    # gets the far form's value bit for bit; a near pair the reference's
    # minimum image (same operations as the reference, so bit-identical when
    # the far images vanish) plus the nine far images from the
    # per-dislocation factors, within 4e-15 of the reference's term relative
    # to |tau| + 1, down to the Taylor branch (r^2 <= 1e-6) and the exact
    # zero; odd under swapping the two dislocations bit for bit
    mc = O.mathcheck()
    rng = np.random.default_rng(14)
    n = 400_000
    xi, yi = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)
    r = 10.0 ** rng.uniform(-5, np.log10(0.3), n)
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    ddx, ddy = r * np.cos(th), r * np.sin(th)
    ddx[: n // 16] = 0.0
    ddy[n // 16: n // 8] = 0.0
    xj, yj = xi - ddx, yi - ddy
    xi[: n // 4] += rng.integers(-3, 4, n // 4)
    yj[n // 4: n // 2] += rng.integers(-3, 4, n // 4)
    out, swp, far_ref = np.empty(n), np.empty(n), np.empty(n)
    isfar, isfar_sw = np.empty(n, np.int32), np.empty(n, np.int32)
    mc.mc_batch_pair_mixed(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(out), O.ip(isfar))
    mc.mc_batch_pair_mixed(n, O.dp(xj), O.dp(yj), O.dp(xi), O.dp(yi), O.dp(swp), O.ip(isfar_sw))
    mc.mc_batch_pair_far(n, O.dp(xi), O.dp(yi), O.dp(xj), O.dp(yj), O.dp(far_ref))
    assert np.array_equal(isfar, isfar_sw)
    far = isfar.astype(bool)
    assert 0.05 < far.mean() < 0.8
    assert np.array_equal(out[far].view(np.int64), far_ref[far].view(np.int64))
    dx, dy = np.empty(n), np.empty(n)
    ex, ey = xi - xj, yi - yj
    mc.mc_batch1(6, n, O.dp(ex), O.dp(dx))
    mc.mc_batch1(6, n, O.dp(ey), O.dp(dy))
    # at dx = 0 exactly (here: x shifted by whole periods) the factors' dx
    # is a rounding residue of either sign, as in the far form
    nz = (out != 0.0) & (dx != 0.0)
    assert np.array_equal(out[nz].view(np.int64), (-swp[nz]).view(np.int64))
    assert np.all(swp[out == 0.0] == 0.0)
    assert np.max(np.abs(out[(dx == 0.0) & ~far])) < 1e-15
    ref, _ = O.xy_batch(dx, dy)
    near = ~far
    assert (near & (dx * dx + dy * dy <= 1e-6)).sum() > 10_000
    err = np.abs(out - ref) / (np.abs(ref) + 1.0)
    assert err[near].max() < 4e-15
    assert err[near].mean() < 2e-16
    # the exact near image dominates: on the close pairs the term is the
    # reference's to a few ulps of itself
    close = near & (np.abs(ref) > 1e3)
    assert close.sum() > 50_000
    assert np.max(np.abs(out[close] - ref[close]) / np.abs(ref[close])) < 1e-15


def _tile(rng, cx, cy, spread, count):
    x = cx + spread * rng.uniform(-1.0, 1.0, 64) + rng.integers(-2, 3, 64)
    y = cy + spread * rng.uniform(-1.0, 1.0, 64) + rng.integers(-2, 3, 64)
    x[count:] = np.nan
    y[count:] = np.nan
    return x, y


def test_tile_kind_bounds_every_pair():
    # a tile is all-far only if every pair's gap 2 sinh^2(pi gx) +
    # 2 sin^2(pi gy) is >= 0.2, all-near only if every one is below, over
    # tiles of random centre, spread and fill, straddling the cell's edges;
    # and the bounds are tight enough that a tile whose extreme distances
    # give a gap beyond 0.25 (below 0.15) is classified all-far (all-near)
    mc = O.mathcheck()
    rng = np.random.default_rng(21)
    kinds = [0, 0, 0]
    for _ in range(6000):
        spread = 10.0 ** rng.uniform(-3.5, -0.7)
        ca, cb = rng.uniform(-0.5, 0.5, 2)
        dist = rng.uniform(0.0, 0.35)
        ang = rng.uniform(0.0, 2.0 * np.pi)
        tx, ty = _tile(rng, ca, cb, spread, int(rng.integers(1, 65)))
        sx, sy = _tile(rng, ca + dist * np.cos(ang), cb + dist * np.sin(ang), spread, int(rng.integers(1, 65)))
        k = mc.mc_tile_kind(O.dp(tx), O.dp(ty), O.dp(sx), O.dp(sy))
        kinds[k] += 1
        ok_t, ok_s = ~np.isnan(tx), ~np.isnan(sx)
        gx = np.abs(tx[ok_t][:, None] - sx[ok_s][None, :])
        gy = np.abs(ty[ok_t][:, None] - sy[ok_s][None, :])
        gx, gy = np.abs(gx - np.rint(gx)), np.abs(gy - np.rint(gy))
        g = 2.0 * np.sinh(np.pi * gx) ** 2 + 2.0 * np.sin(np.pi * gy) ** 2
        if k == 1:
            assert g.min() >= 0.2
        elif k == 0:
            assert g.max() < 0.2
        # the gap at the extremes of gx and gy over the tile (the tile's
        # bounding boxes, which the bounds see)
        lo = 2.0 * np.sinh(np.pi * gx.min()) ** 2 + 2.0 * np.sin(np.pi * gy.min()) ** 2
        hi = 2.0 * np.sinh(np.pi * gx.max()) ** 2 + 2.0 * np.sin(np.pi * gy.max()) ** 2
        if lo >= 0.25:
            assert k == 1
        if hi <= 0.15:
            assert k == 0
    assert min(kinds) > 300


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
