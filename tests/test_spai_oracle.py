"""The CPU restatement of the learned preconditioner (oracle/spai_gnn.py):
its backward pass against central finite differences of its own loss, the
probe hash against SplitMix64's published first output, the pair-block and
first-order terms against dense linear algebra."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import spai_gnn as G  # noqa: E402

import oracle_lib as O  # noqa: E402
from sdddst_amd import configs  # noqa: E402


def _system(n=256, m=2.0, seed=5, h=3e-5):
    x, y, b = configs.dislocations(n, seed)
    Ap, Ai, Ax, dv, diag, raw = O.jacobian(x, y, b, m, h)
    return G.system_from_jacobian(x, y, Ap, Ai, Ax, diag, m / np.sqrt(n)), (Ap, Ai, Ax, diag)


def _dense(Ap, Ai, Ax):
    n = len(Ap) - 1
    A = np.zeros((n, n))
    for j in range(n):
        for k in range(Ap[j], Ap[j + 1]):
            A[Ai[k], j] = Ax[k]
    return A


def test_splitmix_known_answer():
    # SplitMix64 with state 0: the first output is 0xE220A8397B1DCDAF
    z = G.splitmix_mix(np.array([0x9E3779B97F4A7C15], dtype=np.uint64))
    assert int(z[0]) == 0xE220A8397B1DCDAF
    W = G.probes(0, 1, 1)
    assert W[0, 0] == -1.0   # top bit of 0xE220... is set


def test_layout_and_init():
    assert G.NPARAM == 6822
    th = G.init_params(7)
    P = G.unpack(th)
    assert np.all(P["wa"] == 0) and P["beta"][0] == 0 and np.all(P["b0"] == 0)
    assert np.max(np.abs(P["U1"])) <= 1 / np.sqrt(32) and np.max(np.abs(P["W0"])) <= 0.5
    assert np.array_equal(G.pack(P), th)
    assert np.array_equal(th.astype(np.float32).astype(np.float64), th)


def test_csr_view_is_transpose():
    s, (Ap, Ai, Ax, diag) = _system(n=128)
    A = _dense(Ap, Ai, Ax)
    rows = np.repeat(np.arange(s.n), np.diff(Ap))
    assert np.array_equal(s.AxRow, A[rows, Ai])      # row i of A on column i's indices


def test_pair_block_inverts_2x2_blocks():
    s, (Ap, Ai, Ax, diag) = _system(n=256)
    A = _dense(Ap, Ai, Ax)
    paired = np.nonzero(s.bcol >= 0)[0]
    assert len(paired) > 0
    for i in paired[:50]:
        j = s.bcol[i]
        assert s.bcol[j] == i
        inv = np.linalg.inv(A[np.ix_([i, j], [i, j])])
        assert np.allclose([s.bd[i], s.bo[i]], inv[0], rtol=1e-12, atol=0)
    lone = np.nonzero(s.bcol < 0)[0]
    assert np.array_equal(s.bd[lone], 1.0 / np.diag(A)[lone])


def test_core_graph_and_first_order_scales():
    n, m = 256, 2.0
    x, y, b = configs.dislocations(n, 5)
    s, (Ap, Ai, Ax, diag) = _system(n=n, m=m)
    A = _dense(Ap, Ai, Ax)
    rc = m / np.sqrt(n)
    # brute-force core set from the positions
    want = set()
    for i in range(n):
        for j in range(n):
            if i == j or s.bcol[i] == j or A[i, j] == 0 and A[j, i] == 0:
                continue
            lo, hi = min(i, j), max(i, j)
            dx = G.wrap(x[hi] - x[lo])
            dy = G.wrap(y[hi] - y[lo])
            if not (dx * dx + dy * dy > rc * rc):
                want.add((i, j))
    got = set(zip(s.ri.tolist(), s.ci.tolist()))
    assert got == want
    assert np.array_equal(s.ri[s.mirror], s.ci) and np.array_equal(s.mirror[s.mirror], np.arange(len(s.k)))
    # C with gamma = 1 on every edge is -B^-1 E B^-1 (diagonal blocks of
    # B^-1) restricted to the core, each term damped by 1 + |x| + |w|
    bd = s.bd
    x_ = bd[s.ri] * A[s.ri, s.ci]
    w_ = bd[s.ci] * A[s.ci, s.ri]
    assert np.allclose(s.scale_e, -bd[s.ri] * A[s.ri, s.ci] * bd[s.ci] / (1 + np.abs(x_) + np.abs(w_)),
                       rtol=1e-15, atol=0)
    weak = (np.abs(x_) < 1e-3) & (np.abs(w_) < 1e-3)
    assert weak.any()
    assert np.allclose(s.scale_e[weak], (-bd[s.ri] * A[s.ri, s.ci] * bd[s.ci])[weak], rtol=3e-3, atol=0)
    t = np.zeros(n)
    for e in range(len(s.k)):
        i, j = s.ri[e], s.ci[e]
        xw = (bd[i] * A[i, j]) * (bd[j] * A[j, i])
        t[i] += xw / (1 + abs(xw))
    assert np.allclose(s.scale_d, bd * t, rtol=1e-12, atol=1e-300)


def test_zero_readout_is_the_pair_block_method():
    s, (Ap, Ai, Ax, diag) = _system(n=256)
    A = _dense(Ap, Ai, Ax)
    th = G.init_params(3)
    L, g, ge, gd = G.loss_and_grad(th, s, 4, 11)
    assert np.all(ge == 0) and np.all(gd == 0)
    Binv = np.diag(s.bd.copy())
    p = s.bcol >= 0
    Binv[np.nonzero(p)[0], s.bcol[p]] = s.bo[p]
    W = G.probes(11, 4, s.n)
    R = A @ Binv @ W.T - W.T
    q = (R * R).mean(1)
    assert abs(L - np.log1p(q / G.EPS_LOSS).mean()) <= 1e-12 * abs(L)


def test_backward_matches_finite_differences():
    s, _ = _system(n=192, seed=8)
    rng = np.random.default_rng(4)
    th = G.init_params(7) + 0.05 * rng.standard_normal(G.NPARAM)
    L, g, *_ = G.loss_and_grad(th, s, 4, 21)
    P = G.unpack(np.arange(G.NPARAM, dtype=np.float64))
    # indices from every parameter block, chosen where the gradient is nonzero
    checked = 0
    for name, _shape in G.LAYOUT:
        idx = P[name].reshape(-1).astype(int)
        cand = idx[np.abs(g[idx]) > 1e-7 * np.abs(g).max()]
        for t in cand[:2]:
            e = 1e-6 * max(1.0, abs(th[t]))
            tp, tm = th.copy(), th.copy()
            tp[t] += e
            tm[t] -= e
            fd = (G.loss_and_grad(tp, s, 4, 21, with_grad=False)[0] -
                  G.loss_and_grad(tm, s, 4, 21, with_grad=False)[0]) / (2 * e)
            assert abs(fd - g[t]) <= 1e-4 * abs(g[t]) + 1e-9 * np.abs(g).max(), (name, t, fd, g[t])
            checked += 1
    assert checked >= 20
