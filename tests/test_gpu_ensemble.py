"""Batched ensembles against independent runs of the oracle's stepping
driver: each realization's accept/reject sequence, time and positions, with
uneven sizes, zero and constant stress, graph replay against direct
launches, and the regrowth of an overflowing Jacobian capacity."""
import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import configs
from sdddst_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu


def _oracle_runs(sizes, xs, ys, bs, steps, m, sigma):
    out = []
    for n, x, y, b in zip(sizes, xs, ys, bs):
        o = O.OracleSim(x, y, b, cutoff_multiplier=m, prec=1e-6, h0=1e-6,
                        stress="constant" if sigma else "zero", p1=sigma)
        while o.succesful_steps < steps:
            o.step()
        out.append((o.time, o.step_size, o.succesful_steps, o.failed_steps, o.x()))
        o.close()
    return out


def _workload(sizes, seed0):
    parts = [configs.dislocations(n, seed0 + r) for r, n in enumerate(sizes)]
    return [p[0] for p in parts], [p[1] for p in parts], [p[2] for p in parts]


@pytest.mark.parametrize("sigma,mode", [(0.0, "exact"), (0.02, "exact"), (0.0, "fast")])
def test_ensemble_matches_independent_oracle_runs(sigma, mode):
    sizes = [64, 97, 150, 203, 256, 301]
    steps = 15
    xs, ys, bs = _workload(sizes, 3000)
    with Ensemble(sizes, np.concatenate(xs), np.concatenate(ys), np.concatenate(bs), cutoff_multiplier=1.0,
                  precision=1e-6, field_mode=mode, sigma=sigma) as e:
        attempts = e.run(steps, max_attempts=40 * steps)
        t, h, s, f = e.state()
        xg = e.split(e.positions())
        st = e.stats()
    ref = _oracle_runs(sizes, xs, ys, bs, steps, 1.0, sigma)
    assert attempts == max(r[2] + r[3] for r in ref)
    for k, (tr, hr, sr, fr, xr) in enumerate(ref):
        assert s[k] == sr and f[k] == fr, (k, s[k], sr, f[k], fr)
        assert abs(t[k] - tr) <= 1e-9 * abs(tr)
        assert abs(h[k] - hr) <= 1e-6 * abs(hr)
        assert np.max(np.abs(xg[k] - xr)) < 1e-9
    assert st.solves > 0 and st.lu_fallbacks == 0


def test_graph_replay_equals_direct_launches():
    sizes = configs.ensemble_sizes(12, 3, 64, 400)
    xs, ys, bs = _workload(sizes, 3000)
    res = []
    for graph in (True, False):
        with Ensemble(sizes, np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)) as e:
            e.run(10, max_attempts=400, use_graph=graph)
            res.append((e.state(), e.positions()))
    (a, xa), (b, xb) = res
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(xa.view(np.int64), xb.view(np.int64))


def test_autotune_and_capacity_regrowth():
    # one realization clustered in a small square: every pair lies inside
    # the predicate radius, more entries than the default capacity
    sizes = [80, 520]
    x0, y0, b0 = configs.dislocations(80, 11)
    x1, y1, b1 = configs.dislocations(520, 12)
    x1 = x1 * 0.02
    y1 = y1 * 0.02
    with Ensemble(sizes, np.concatenate([x0, x1]), np.concatenate([y0, y1]), np.concatenate([b0, b1]),
                  field_mode="exact") as e:
        report = e.autotune()
        assert "chosen" in report
        st0 = e.stats()
        assert st0.regrows >= 1
        e.run(3, max_attempts=200)
        t, h, s, f = e.state()
        xg = e.split(e.positions())
        st = e.stats()
    ref = _oracle_runs(sizes, [x0, x1], [y0, y1], [b0, b1], 3, 1.0, 0.0)
    for k, (tr, hr, sr, fr, xr) in enumerate(ref):
        assert s[k] == sr and f[k] == fr
        assert abs(t[k] - tr) <= 1e-9 * abs(tr)
        assert np.max(np.abs(xg[k] - xr)) < 1e-9
    assert st.speeds_block in (0, 64, 128, 256) and st.solver_block in (256, 512, 1024)
    assert st.spmv_lanes in (4, 8, 16, 64)
