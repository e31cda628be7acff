"""EV analysis (spectral_decompositor.cpp:63-201, time_series_processor.cpp:
40-111, 205-246): the device Hessian, eigen-decomposition, per-mode
observables and eigenvalue derivatives against a numpy restatement on the
oracle's field derivatives and point-defect term, with numpy's symmetric
eigensolver."""
import numpy as np
import pytest

import oracle_lib as O
import sdddst_amd as S
from sdddst_amd import configs


def _normalize(v):
    lib = O.oracle()
    return np.array([lib.orc_normalize(float(t)) for t in v])


def oracle_hessian(x, y, b, px=None, py=None):
    """The reference's Hessian: b_i b_j dtau/dx for i > j (mirrored), the
    diagonal minus the lower row, minus the point-defect terms (A and K of
    spectral_decompositor.cpp:81-90), minus the column below, in that order."""
    n = len(x)
    I, J = np.tril_indices(n, -1)
    dx, dy = _normalize(x[I] - x[J]), _normalize(y[I] - y[J])
    _, xyd = O.xy_batch(dx, dy)
    t = b[I] * b[J] * xyd
    H = np.zeros((n, n))
    H[I, J] = t
    H[J, I] = t
    pc = 0 if px is None else len(px)
    A = 1e-4 * 16.0 / np.sqrt(pc) if pc else 0.0
    K = 1.65 * 1.65 * 1e6 / 256.0 * n if pc else 0.0
    lib = O.oracle()
    for i in range(n):
        sub = 0.0
        for j in range(i):
            sub += H[i, j]
        d = 0.0 - sub
        for k in range(pc):
            d -= lib.orc_pd_diag_term(lib.orc_normalize(float(x[i] - px[k])), lib.orc_normalize(float(y[i] - py[k])),
                                      float(b[i]), A, K, 1.0)
        for j in range(i + 1, n):
            d -= H[j, i]
        H[i, i] = d
    return H


def _x2_matrix(x, y):
    """X_ij = xy_diff_x2(wrap(x_i - x_j), wrap(y_i - y_j)) for i != j."""
    lib = O.oracle()
    n = len(x)
    X = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                X[i, j] = lib.orc_xy_diff_x2(lib.orc_normalize(float(x[i] - x[j])),
                                             lib.orc_normalize(float(y[i] - y[j])))
    return X


def oracle_deriv_matrix(x, y, b, s):
    """C_ij = sum_k s_k T_ijk over calculateHessianDerivative's tensor
    (time_series_processor.cpp:205-246), reduced: (s_i - s_j) b_i b_j X_ij
    off the diagonal, minus the row's sum on it."""
    C = (s[:, None] - s[None, :]) * np.outer(b, b) * _x2_matrix(x, y)
    np.fill_diagonal(C, 0.0)
    np.fill_diagonal(C, -C.sum(axis=1))
    return C


def test_deriv_matrix_reduces_the_reference_tensor():
    # the reference's per-mode sum over i, j, k of v_i v_j s_k T_ijk / sigma',
    # with T built term by term as calculateHessianDerivative builds it,
    # against the reduced matrix
    n = 9
    x, y, b = configs.dislocations(n, 5)
    s = np.random.default_rng(6).uniform(-1, 1, n)
    X = _x2_matrix(x, y)
    T = np.zeros((n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                if i == j == k:
                    T[i, j, k] = -b[i] * sum(b[m] * X[i, m] for m in range(n) if m != i)
                elif i == j:
                    T[i, j, k] = b[i] * b[k] * X[i, k]
                elif i == k:
                    T[i, j, k] = b[i] * b[j] * X[i, j]
                elif j == k:
                    T[i, j, k] = -b[i] * b[j] * X[i, j]
    C = oracle_deriv_matrix(x, y, b, s)
    np.testing.assert_allclose(np.einsum("ijk,k->ij", T, s), C, rtol=1e-12, atol=1e-12 * np.abs(C).max())
    _, V = np.linalg.eigh(oracle_hessian(x, y, b))
    rate = 0.7
    ref = np.array([sum(V[i, m] * V[j, m] / rate * (T[i, j] @ s) for i in range(n) for j in range(n))
                    for m in range(n)])
    np.testing.assert_allclose(np.einsum("im,ij,jm->m", V, C, V) / rate, ref, rtol=1e-10,
                               atol=1e-12 * np.abs(ref).max())


def test_deriv_is_the_eigenvalue_slope_along_the_speeds():
    # C is the derivative of the Hessian (no point defects) along x + e s, so
    # v^T C v is the slope of the eigenvalue; the reference's xy_diff_x2
    # lacks the field's factor 2 pi^2 (AnalyticField.cpp:95-106), which the
    # slope carries
    n = 24
    x, y, b = configs.dislocations(n, 7)
    s = np.random.default_rng(8).uniform(-1, 1, n)
    lam, V = np.linalg.eigh(oracle_hessian(x, y, b))
    d = np.einsum("im,ij,jm->m", V, oracle_deriv_matrix(x, y, b, s), V) * 2.0 * np.pi ** 2
    e = 1e-6
    lp = np.linalg.eigvalsh(oracle_hessian(x + e * s, y, b))
    lm = np.linalg.eigvalsh(oracle_hessian(x - e * s, y, b))
    fd = (lp - lm) / (2 * e)
    scale = np.abs(lam).max()
    gaps = np.minimum(np.r_[np.inf, np.diff(lam)], np.r_[np.diff(lam), np.inf])
    sep = gaps > 1e-3 * scale
    assert sep.sum() > n // 2
    assert np.abs(fd - d)[sep].max() <= 1e-5 * np.abs(d).max()


def test_oracle_hessian_is_symmetric_and_translation_invariant():
    # without point defects every row sums to zero (moving all dislocations
    # together changes no force): H 1 = 0, so the uniform glide mode has
    # eigenvalue 0
    x, y, b = configs.dislocations(96, 12)
    H = oracle_hessian(x, y, b)
    assert np.array_equal(H, H.T)
    scale = np.abs(H).max()
    assert np.abs(H.sum(axis=1)).max() <= 1e-12 * scale * len(x)
    lam, V = np.linalg.eigh(H)
    k = np.argmin(np.abs(lam))
    assert abs(lam[k]) <= 1e-10 * np.abs(lam).max()
    assert abs(V[:, k].sum()) / np.sqrt(len(x)) > 1 - 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("pc", [0, 24])
def test_ev_decompose_matches_oracle(pc):
    n = 192
    x, y, b = configs.dislocations(n, 31)
    kw, px, py = {}, None, None
    if pc:
        px, py = configs.point_defects(n, pc, 31)
        kw = dict(pc=pc)
    H = oracle_hessian(x, y, b, px, py)
    lam_r, V_r = np.linalg.eigh(H)
    with S.Context(n, **kw) as c:
        c.upload_state(0, x, y, b)
        if pc:
            c.upload_points(px, py)
        out = c.ev_decompose(0)
    lam = out["eigenvalues"]
    scale = np.abs(lam_r).max()
    assert np.all(np.diff(lam) >= 0)
    assert np.abs(lam - lam_r).max() <= 1e-11 * scale
    # eigenvectors (up to sign) of the modes separated from their neighbours
    # by more than 1e-8 of the spectrum's scale: with the two solvers' errors
    # near 1e-13 of it the angle is below 1e-5
    gaps = np.minimum(np.r_[np.inf, np.diff(lam_r)], np.r_[np.diff(lam_r), np.inf])
    sep = gaps > 1e-8 * scale
    assert sep.sum() > n // 2
    V = out["eigenvectors"]
    dots = np.abs(np.einsum("ij,ji->i", V, V_r))
    assert np.all(dots[sep] > 1 - 1e-9)
    # the observables of the device's own eigenpairs (the reduction kernel)
    np.testing.assert_allclose(out["pn"], 1.0 / np.sum(V ** 4, axis=1), rtol=1e-12)
    s = V @ b
    np.testing.assert_allclose(out["sum_b"], s, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(out["gan"], s ** 2 / lam, rtol=1e-9, atol=1e-12 * np.abs(s ** 2 / lam).max())
    if pc == 0:
        # the uniform glide mode
        k = np.argmin(np.abs(lam))
        assert abs(lam[k]) <= 1e-10 * scale
        assert abs(V[k].sum()) / np.sqrt(n) > 1 - 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("pc", [0, 24])
def test_ev_derivatives_match_oracle(pc):
    # the device's C (xy_diff_x2 per ordered pair), dgemm and per-mode dots
    # against the oracle's C on the device's own eigenvectors
    n, rate = 128, 2.5
    x, y, b = configs.dislocations(n, 33)
    s = np.random.default_rng(34).uniform(-1, 1, n)
    kw = dict(pc=pc) if pc else {}
    with S.Context(n, **kw) as c:
        c.upload_state(0, x, y, b)
        if pc:
            px, py = configs.point_defects(n, pc, 33)
            c.upload_points(px, py)
        plain = c.ev_decompose(0)
        out = c.ev_decompose(0, speeds=s, stress_rate=rate)
    for k in ("eigenvalues", "eigenvectors", "pn", "gan", "sum_b"):
        assert np.array_equal(out[k], plain[k])
    V = out["eigenvectors"]
    C = oracle_deriv_matrix(x, y, b, s)
    want = np.einsum("mi,ij,mj->m", V, C, V) / rate
    atol = 1e-12 * np.abs(C).sum(axis=1).max() / rate
    np.testing.assert_allclose(out["deriv"], want, rtol=1e-10, atol=atol)


@pytest.mark.gpu
def test_ev_analysis_rejects_speeds_without_outputs():
    # speeds and deriv go together; a lone one is an argument error, not a
    # silent decomposition without derivatives
    import ctypes
    n = 32
    x, y, b = configs.dislocations(n, 35)
    with S.Context(n) as c:
        c.upload_state(0, x, y, b)
        lam = np.empty(n)
        sp = np.zeros(n)
        dp = ctypes.POINTER(ctypes.c_double)
        rc = c.lib.sdddst_ev_analysis(c.h, 0, sp.ctypes.data_as(dp), 1.0, lam.ctypes.data_as(dp), None, None, None,
                                      None, None)
        assert rc != 0 and b"speeds and deriv go together" in c.lib.sdddst_last_error()
        out = c.ev_decompose(0, vectors=False)
        assert np.all(np.diff(out["eigenvalues"]) >= 0)
