"""EV analysis (spectral_decompositor.cpp:63-201, time_series_processor.cpp:
40-82): the device Hessian, eigen-decomposition and per-mode observables
against a numpy restatement on the oracle's field derivative and point-defect
term, with numpy's symmetric eigensolver."""
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
