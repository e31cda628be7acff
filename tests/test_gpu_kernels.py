"""GPU parity of the per-pair field, the device elementary functions, the
pair-speed kernel (K1) and the Jacobian assembly (K2) against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
import sdddst_amd as S
from sdddst_amd import configs

pytestmark = pytest.mark.gpu


def _ctx(n, **kw):
    return S.Context(n, **kw)


@pytest.fixture(scope="module")
def ctx256():
    c = _ctx(256)
    yield c
    c.close()


# -- device elementary functions -------------------------------------------

@pytest.mark.parametrize("fn,lo,hi", [(0, -3.1416, 3.1416), (0, -0.02, 0.02), (1, -3.1416, 3.1416), (2, -40, 40),
                                      (3, -35, 35), (3, -0.35, 0.35), (4, -35, 35), (4, -0.35, 0.35),
                                      (5, -35, 35), (6, 0, 4), (8, -3, 3)])
def test_device_libm_matches_host_bits(ctx256, fn, lo, hi):
    rng = np.random.default_rng(100 + fn)
    x = rng.uniform(lo, hi, 400_000)
    dev = ctx256.libm_eval(fn, x)
    host = O.libm_batch(fn, x)
    mism = np.count_nonzero(dev.view(np.int64) != host.view(np.int64))
    # sin(2 pi d) of the point-defect path: one measured non-reproduced case
    # in 3e5 (TAYLOR_SIN region); everything else is bit-exact.
    allowed = 2 if fn == 1 else 0
    assert mism <= allowed, f"fn {fn}: {mism} bit mismatches"


def test_device_pow_one_third(ctx256):
    rng = np.random.default_rng(7)
    t = np.concatenate([rng.uniform(0.05, 100, 200_000), np.exp(rng.uniform(-20, 20, 100_000))])
    dev = ctx256.libm_eval(7, t)
    host = O.libm_batch(7, t)
    ulps = np.abs(dev.view(np.int64) - host.view(np.int64))
    assert ulps.max() <= 1
    assert np.count_nonzero(ulps) <= 0.005 * len(t)


# -- per-pair field --------------------------------------------------------

def _pairs(n, seed):
    rng = np.random.default_rng(seed)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, n), rng.uniform(-2e-3, 2e-3, n // 4),
                         np.array([0.0, -0.5, 0.49999999999999994, 1e-9, -1e-9])])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, n), rng.uniform(-2e-3, 2e-3, n // 4),
                         np.array([0.3, 0.1, -0.2, 2e-4, -3e-4])])
    return dx, dy


def test_field_xy_exact_is_bit_identical(ctx256):
    dx, dy = _pairs(200_000, 1)
    dev = ctx256.field_eval(0, dx, dy)
    ref, _ = O.xy_batch(dx, dy)
    assert np.array_equal(dev.view(np.int64), ref.view(np.int64))


def test_field_xy_diff_x(ctx256):
    dx, dy = _pairs(100_000, 2)
    dev = ctx256.field_eval(1, dx, dy)
    _, ref = O.xy_batch(dx, dy)
    same = dev.view(np.int64) == ref.view(np.int64)
    assert same.mean() > 0.998
    rel = np.abs(dev - ref) / np.maximum(np.abs(ref), 1.0)
    assert rel.max() < 1e-13


def test_field_xy_fast_abs_error(ctx256):
    dx, dy = _pairs(200_000, 3)
    dev = ctx256.field_eval(2, dx, dy)
    ref, _ = O.xy_batch(dx, dy)
    err = np.abs(dev - ref) / (np.abs(ref) + 1.0)
    assert err.max() < 4e-15


def test_field_xy_diff_x2_golden(ctx256):
    # AnalyticField::xy_diff_x2 (AnalyticField.cpp:95-132, 183-202) against
    # the golden vectors made from the reference's own translation unit; below
    # r = 0.05 the reference's numerator (no small-r branch) magnifies one-ulp
    # differences of pow(x, 3.0) against a correctly rounded cube
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "field_points_dx2.npz"))
    dev = ctx256.field_eval(3, g["dx"], g["dy"])
    ref = g["xy_diff_x2"]
    r = np.hypot(g["dx"], g["dy"])
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(dev), fin)
    rel = np.abs(dev[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1.0)
    far = r[fin] > 0.05
    assert far.sum() > 1000
    assert rel[far].max() < 1e-13
    assert rel.max() < 1e-8
    assert np.mean(dev[fin].view(np.int64) == ref[fin].view(np.int64)) > 0.99


def test_analytic_field_class_methods(ctx256):
    # the drop-in's AnalyticField::xy / xy_diff_x / xy_diff_x2 (Field.h:34-36)
    # evaluate the device field: same bits as the array evaluation
    from sdddst_amd.simulation import analytic_field
    dx, dy = _pairs(2000, 4)
    for which, dev_which in ((0, 0), (1, 1), (2, 3)):
        a = analytic_field(which, dx, dy)
        d = ctx256.field_eval(dev_which, dx, dy)
        assert np.array_equal(a.view(np.int64), d.view(np.int64))


# -- K1 pair speeds --------------------------------------------------------

def _speeds_dev(n, seed, mode, sigma=0.0, cutoff=1.0):
    x, y, b = configs.dislocations(n, seed)
    with _ctx(n, field_mode=mode, cutoff_multiplier=cutoff) as c:
        c.upload_state(0, x, y, b)
        v = c.speeds(0, S.V_INIT, sigma, update_tol=True)
        tol = c.tolerance()
    return x, y, b, v, tol


def test_speeds_ordered_bit_identical_to_oracle():
    x, y, b, v, tol = _speeds_dev(256, 1, S.FIELD_ORDERED)
    vr, tr = O.speeds(x, y, b)
    assert np.array_equal(v.view(np.int64), vr.view(np.int64))
    assert np.array_equal(tol.view(np.int64), tr.view(np.int64))


def test_speeds_ordered_bit_identical_reference_loop():
    ref = O.reference()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    x, y, b, v, _ = _speeds_dev(512, 11, S.FIELD_ORDERED, sigma=0.02)
    vr = np.empty(512)
    ref.ref_speeds(512, O.dp(x), O.dp(y), O.dp(b), 0.02, O.dp(vr))
    assert np.array_equal(v.view(np.int64), vr.view(np.int64))


def _speed_parity(v, vr, x, y, b):
    # per-component 1e-12 relative; a conditioning-aware bound for the
    # components whose sum cancels: |dv| <= 1e-12 * sum_j |t_ij| scale
    rel = np.abs(v - vr) / np.abs(vr)
    return rel


@pytest.mark.parametrize("mode", [S.FIELD_EXACT, S.FIELD_FAST])
@pytest.mark.parametrize("n,seed", [(256, 1), (1000, 5), (2048, 9)])
def test_speeds_tiled_vs_oracle(mode, n, seed):
    x, y, b, v, tol = _speeds_dev(n, seed, mode, sigma=0.02)
    vr, tr = O.speeds(x, y, b, sigma=0.02)
    assert np.array_equal(tol.view(np.int64), tr.view(np.int64))
    rel = np.abs(v - vr) / np.abs(vr)
    assert np.median(rel) < 1e-14
    strict, ok = O.speed_parity(v, vr, O.speeds_abs(x, y, b), n)
    assert ok.all()
    assert strict.mean() >= 0.995
    assert np.linalg.norm(v - vr) / np.linalg.norm(vr) < 1e-13


# Sizes where most tiles are x-separated and summed by the field's Fourier
# series in y (the fast field's far-tile path); 16384 is the headline system.
@pytest.mark.parametrize("n,seed", [(8192, 13), (16384, 2)])
def test_speeds_fast_series_vs_oracle(n, seed):
    x, y, b, v, tol = _speeds_dev(n, seed, S.FIELD_FAST, sigma=0.02)
    vr, tr = O.speeds(x, y, b, sigma=0.02)
    assert np.array_equal(tol.view(np.int64), tr.view(np.int64))
    strict, ok = O.speed_parity(v, vr, O.speeds_abs(x, y, b), n)
    assert ok.all()
    assert strict.mean() >= 0.995
    assert np.linalg.norm(v - vr) / np.linalg.norm(vr) < 1e-13


def test_speeds_sum_rule_and_determinism_n16384():
    n = 16384
    x, y, b = configs.dislocations(n, 2)
    with _ctx(n, field_mode=S.FIELD_FAST, cutoff_multiplier=2.0) as c:
        c.upload_state(0, x, y, b)
        v1 = c.speeds(0, S.V_INIT, 0.02)
        v2 = c.speeds(0, S.V_INIT2, 0.02)
    assert np.array_equal(v1.view(np.int64), v2.view(np.int64))
    # pair terms cancel exactly in exact arithmetic: sum v = sigma * sum b = 0
    assert abs(v1.sum() - 0.02 * b.sum()) < 1e-9 * np.abs(v1).sum()


# -- K2 Jacobian -----------------------------------------------------------

# 20000 columns: the count scan's second, partial tile of 16384
@pytest.mark.parametrize("n,m,seed", [(256, 1.0, 1), (300, 1.0, 4), (1024, 2.0, 7), (64, 1.0, 3), (4096, 1.0, 9),
                                      (4096, 2.0, 10), (4096, 4.0, 11), (20000, 1.0, 12)])
def test_jacobian_pattern_bit_exact(n, m, seed):
    x, y, b = configs.dislocations(n, seed)
    h = 1e-4
    with _ctx(n, cutoff_multiplier=m, field_mode=S.FIELD_EXACT) as c:
        c.upload_state(0, x, y, b)
        Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h)
    rAp, rAi, rAx, rdv, ridx, raw = O.jacobian(x, y, b, m, h)
    assert np.array_equal(Ap, rAp)
    assert np.array_equal(Ai, rAi)
    assert np.array_equal(dg, ridx)
    rel = np.abs(Ax - rAx) / np.maximum(np.abs(rAx), 1e-300)
    assert np.mean(Ax.view(np.int64) == rAx.view(np.int64)) > 0.995
    assert rel.max() < 1e-12
    assert np.max(np.abs(dv - rdv)) < 1e-13


def test_jacobian_full_size_n16384_m2():
    # the headline workload's matrix (N = 16384, m = 2, seed 2): pattern,
    # diagonal positions and values against the oracle's column-by-column
    # assembly (simulation.cpp:236-384), exact field
    n, m, h = 16384, 2.0, 1e-4
    x, y, b = configs.dislocations(n, 2)
    with _ctx(n, cutoff_multiplier=m, field_mode=S.FIELD_EXACT) as c:
        c.upload_state(0, x, y, b)
        Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h)
    rAp, rAi, rAx, rdv, ridx, raw = O.jacobian(x, y, b, m, h, cap=n * 700)
    assert rAp[-1] > 1.0e7
    assert np.array_equal(Ap, rAp)
    assert np.array_equal(Ai, rAi)
    assert np.array_equal(dg, ridx)
    assert np.mean(Ax.view(np.int64) == rAx.view(np.int64)) > 0.995
    assert np.max(np.abs(Ax - rAx) / np.maximum(np.abs(rAx), 1e-300)) < 1e-12
    assert np.max(np.abs(dv - rdv)) < 1e-13


@pytest.mark.parametrize("n,m,seed", [(300, 1.0, 4), (1024, 2.0, 7)])
def test_jacobian_fast_field(n, m, seed):
    # production derivative: same pattern, values within 1e-12 of max(|A|, h)
    x, y, b = configs.dislocations(n, seed)
    h = 1e-4
    with _ctx(n, cutoff_multiplier=m, field_mode=S.FIELD_FAST) as c:
        c.upload_state(0, x, y, b)
        Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h)
    rAp, rAi, rAx, rdv, ridx, raw = O.jacobian(x, y, b, m, h)
    assert np.array_equal(Ap, rAp) and np.array_equal(Ai, rAi) and np.array_equal(dg, ridx)
    assert np.max(np.abs(Ax - rAx) / np.maximum(np.abs(rAx), h)) < 1e-12
    assert np.max(np.abs(dv - rdv)) < 1e-13


def _assemble(c, h, rebuild):
    """Assemble with the given rebuild mode (1: full, 2: reuse the pattern of
    the last build) and download the CSC, d and the diagonal positions."""
    import ctypes
    from sdddst_amd._capi import _dptr, _iptr
    nnz = ctypes.c_int64()
    c._check(c.lib.sdddst_jacobian_csc(c.h, 0, h, rebuild, None, None, None, None, None, ctypes.byref(nnz)))
    Ap = np.empty(c.n + 1, dtype=np.int32)
    Ai = np.empty(nnz.value, dtype=np.int32)
    Ax = np.empty(nnz.value)
    dv = np.empty(c.n)
    dg = np.empty(c.n, dtype=np.int32)
    c._check(c.lib.sdddst_jacobian_csc(c.h, 0, h, 0, _iptr(Ap), _iptr(Ai), _dptr(Ax), _dptr(dv), _iptr(dg),
                                       ctypes.byref(nnz)))
    return Ap, Ai, Ax, dv, dg


def test_assembly_outgrowing_its_storage_assembles_again():
    # assemblies after the first run against the storage the first one sized
    # and read the entry count afterwards; a configuration with far more
    # neighbours overflows it: the library grows the storage and assembles
    # again, with the same matrix as a fresh context
    n, m, h = 2048, 1.0, 3e-5
    x, y, b = configs.dislocations(n, 77)
    xs, ys = 0.25 * x, 0.25 * y          # 16x the density: about 16x the entries
    with S.Context(n, cutoff_multiplier=m, field_mode=S.FIELD_FAST) as c:
        c.upload_state(0, x, y, b)
        small = _assemble(c, h, 1)
        c.upload_state(0, xs, ys, b)
        grown = _assemble(c, h, 1)
    with S.Context(n, cutoff_multiplier=m, field_mode=S.FIELD_FAST) as c:
        c.upload_state(0, xs, ys, b)
        fresh = _assemble(c, h, 1)
    assert len(grown[1]) > 4 * len(small[1])
    for u, v in zip(grown, fresh):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))


def _pair_solve(c, h):
    c.solver_select(S.SOLVER_FGMRES, 1e-10, 300, 40)
    c.set_preconditioner(S.PRECOND_PAIR)
    c.residual(h, 0, 0, S.V_INIT, S.V_INIT2)
    c.solver_factor()
    return c.solver_solve(download=True)


@pytest.mark.parametrize("pc, ratio", [(0, 0.5), (40, 0.5), (0, 0.3), (40, 1.0)])
def test_stage_two_reuse_is_bit_identical(pc, ratio):
    # stage II reuses stage I's unsigned values times 0.5: the matrix, d and
    # the pair blocks equal a fresh assembly at h / 2 bit for bit; a reuse of
    # the pattern at another step size (0.3 h: the values again from tau_x and
    # the taper; h: the values of the build) equals a fresh assembly as well
    n, m, h = 1024, 2.0, 3e-5
    x, y, b = configs.dislocations(n, 61)
    kw = {}
    if pc:
        px, py = configs.point_defects(n, pc, 61)
        kw = dict(pc=pc, A=1e-4 * 16.0 / np.sqrt(n), KASQR=1.65 * 1.65 * 1e6 / 256.0 * n)
    out = []
    for reuse in (True, False):
        with S.Context(n, cutoff_multiplier=m, field_mode=S.FIELD_FAST, **kw) as c:
            c.upload_state(0, x, y, b)
            if pc:
                c.upload_points(px, py)
            c.speeds(0, S.V_INIT, 0.0)
            c.speeds(0, S.V_INIT2, 0.0)
            if reuse:
                _assemble(c, h, 1)
                csc = _assemble(c, ratio * h, 2)
            else:
                csc = _assemble(c, ratio * h, 1)
            out.append((csc, _pair_solve(c, ratio * h)))
    (a, da), (b_, db) = out
    for u, v in zip(a, b_):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    assert np.array_equal(da.view(np.int64), db.view(np.int64))


# -- point defects (config-5 style; A and KASQR scaled as the CLI does) --------

def _pd_setup(n, pc, seed):
    x, y, b = configs.dislocations(n, seed)
    px, py = configs.point_defects(n, pc, seed)
    A, K = 1e-4 * 16.0 / np.sqrt(n), 1.65 * 1.65 * 1e6 / 256.0 * n
    return x, y, b, px, py, A, K


@pytest.mark.parametrize("mode", [S.FIELD_ORDERED, S.FIELD_EXACT])
def test_speeds_with_point_defects(mode):
    n, pc = 512, 96
    x, y, b, px, py, A, K = _pd_setup(n, pc, 50)
    with S.Context(n, pc=pc, field_mode=mode, A=A, KASQR=K) as c:
        c.upload_state(0, x, y, b)
        c.upload_points(px, py)
        v = c.speeds(0, S.V_INIT, 0.01)
        tol = c.tolerance()
    vr, tr = O.speeds(x, y, b, sigma=0.01, px=px, py=py, A=A, K=K)
    assert np.array_equal(tol.view(np.int64), tr.view(np.int64))
    if mode == S.FIELD_ORDERED:
        assert np.array_equal(v.view(np.int64), vr.view(np.int64))
    else:
        strict, ok = O.speed_parity(v, vr, O.speeds_abs(x, y, b), n)
        assert ok.all() and strict.mean() >= 0.995


def test_jacobian_with_point_defects():
    n, pc, h = 400, 64, 1e-4
    x, y, b, px, py, A, K = _pd_setup(n, pc, 51)
    with S.Context(n, pc=pc, cutoff_multiplier=2.0, A=A, KASQR=K, field_mode=S.FIELD_EXACT) as c:
        c.upload_state(0, x, y, b)
        c.upload_points(px, py)
        Ap, Ai, Ax, dv, dg = c.jacobian_csc(0, h)
    rAp, rAi, rAx, rdv, ridx, raw = O.jacobian(x, y, b, 2.0, h, px=px, py=py, A=A, K=K)
    assert np.array_equal(Ap, rAp) and np.array_equal(Ai, rAi) and np.array_equal(dg, ridx)
    rel = np.abs(Ax - rAx) / np.maximum(np.abs(rAx), 1e-300)
    assert rel.max() < 1e-11
    assert np.max(np.abs(dv - rdv)) < 1e-12


def test_fp64_probe_rates_are_physical():
    # the sustained FP64 stream rates lie below the issue limit of one
    # wave instruction per SIMD every 4 clocks at 2.5 GHz, and the FMA
    # stream's flops below the published 78.6 TFLOP/s (+5 %)
    from sdddst_amd._capi import fp64_probe
    tf, fma_rate, mul_rate = fp64_probe(4000)
    print(f"fp64 probe: {tf:.1f} TFLOP/s, {fma_rate:.3e} / {mul_rate:.3e} VALU/s", flush=True)
    limit = 256 * 4 * 2.5e9 / 4
    assert 0.1 * limit < fma_rate <= limit and 0.1 * limit < mul_rate <= limit
    assert 5.0 < tf <= 78.6 * 1.05
