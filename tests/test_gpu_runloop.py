"""The run loop of the sdddst executable on the device against the oracle's
restatement of Simulation::run and stage III (reference simulation.cpp:
417-433, 553-570, 603-633): the time, strain, step-count, speed, stress and
avalanche limits, the cutoff switch to infinity, the post-relax steps, the
sub-configuration files, the spring protocol (spring_protocol.cpp:33-59),
a forced rejection from a large first step, and the coincident-pair
tolerance path (precision_handler.cpp:77-81).

Each threshold is taken halfway between two consecutive logged values of an
unconstrained oracle run, so the step at which it acts is the same on both
sides although the logged values agree only to ~1e-12."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import sdddst_amd as S
from sdddst_amd import configs
from sdddst_amd.simulation import read_log

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sdddst_amd", "sdddst")
N, SEED, M = 96, 71, 1.0      # m = 1 < sqrt(2N)/12: a finite cutoff


def _setup(tmp_path):
    x, y, b = configs.dislocations(N, SEED)
    d = tmp_path / "in.dconf"
    configs.write_dconf(str(d), x, y, b)
    return configs.read_dconf(str(d)), d


def _oracle(xyb, stress="zero", p1=0.0, p2=0.0, h0=1e-6):
    x, y, b = xyb
    return O.OracleSim(x, y, b, cutoff_multiplier=M, prec=1e-6, h0=h0, stress=stress, p1=p1, p2=p2,
                       calculate_strain=True)


def _mid(log, row, col):
    return 0.5 * (log[row, col] + log[row + 1, col])


def _cli(tmp_path, d, *args):
    out, log = tmp_path / "out.dconf", tmp_path / "run.log"
    cmd = [EXE, "-c", "--dislocation-configuration", str(d), "--result-dislocation-configuration", str(out),
           "--logfile-path", str(log), "--cutoff-multiplier", repr(M), "--position-precision", "1e-6",
           "--calculate-strain", "--field-mode", "exact", "--linear-solver", "dense-lu",
           *[("%.17g" % a) if isinstance(a, (float, np.floating)) else str(a) for a in args]]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return configs.read_dconf(str(out))[0], read_log(str(log)), r


def _compare(o, xg, lg):
    lo = o.log()
    assert lg.shape == lo.shape
    assert np.array_equal(lg[:, 1], lo[:, 1]) and np.array_equal(lg[:, 2], lo[:, 2])   # step counts
    cols = [0, 4, 5, 7, 9, 10, 11]
    a, r = lg[:, cols], lo[:, cols]
    assert (np.isnan(a) == np.isnan(r)).all()
    fin = ~np.isnan(r)
    assert np.max(np.abs(a[fin] - r[fin]) / np.maximum(np.abs(r[fin]), 1e-300)) < 1e-8
    assert np.max(np.abs(xg - o.x())) < 1e-9


def _free_log(xyb, steps, **kw):
    o = _oracle(xyb, **kw)
    o.run(step_count_limit=steps)
    log = o.log()
    o.close()
    return log


def test_speed_limit_and_post_relax(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 14), 10, 4)            # mean |v| falls through thr at step 11
    o = _oracle(xyb)
    o.run(speed_limit=thr, post_relax=3)
    assert o.run_state()["finish"] == 1.0 and o.succesful_steps == 14
    xg, lg, _ = _cli(tmp_path, d, "--speed-limit", thr, "--post-relax", "3")
    _compare(o, xg, lg)
    o.close()


def test_time_limit(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 12), 8, 0)
    o = _oracle(xyb)
    o.run(time_limit=thr)
    assert o.succesful_steps == 9
    xg, lg, _ = _cli(tmp_path, d, "--time-limit", thr)
    _compare(o, xg, lg)
    o.close()


def test_stress_limit_fixed_rate(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 12, stress="fixed-rate", p1=200.0), 9, 7)
    o = _oracle(xyb, stress="fixed-rate", p1=200.0)
    o.run(stress_limit=thr)
    assert o.succesful_steps == 10
    xg, lg, _ = _cli(tmp_path, d, "--fixed-rate-external-stress", "200", "--stress-limit", thr)
    _compare(o, xg, lg)
    o.close()


def test_strain_increase_limit(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 12, stress="constant", p1=40.0), 7, 9)
    o = _oracle(xyb, stress="constant", p1=40.0)
    o.run(strain_limit=thr)
    assert o.succesful_steps == 8
    xg, lg, _ = _cli(tmp_path, d, "--constant-external-stress", "40", "--strain-increase-limit", thr)
    _compare(o, xg, lg)
    o.close()


def test_avalanche_detection_limit(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 10), 5, 4)              # above thr up to step 5, below from step 6
    o = _oracle(xyb)
    o.run(avalanche_limit=1, avalanche_threshold=thr)
    st = o.run_state()
    assert st["avalanche_count"] == 1.0 and st["in_avalanche"] == 0.0 and o.succesful_steps == 6
    xg, lg, _ = _cli(tmp_path, d, "--avalanche-detection-limit", "1", "--avalanche-speed-threshold", thr)
    _compare(o, xg, lg)
    o.close()


def test_cutoff_switch_to_infinity(tmp_path):
    xyb, d = _setup(tmp_path)
    thr = _mid(_free_log(xyb, 12), 6, 4)              # mean |v| < thr from step 7 on
    o = _oracle(xyb)
    o.run(step_count_limit=12, cutoff_inf_threshold=thr)
    lo = o.log()
    rc = 1.0 / np.sqrt(N) * M
    # the row of the switch still logs the old cutoff (simulation_data.cpp:208-218)
    assert lo[7, 5] == rc and lo[8, 5] == 1.0 / np.sqrt(N) * 1e20
    xg, lg, _ = _cli(tmp_path, d, "--step-count-limit", "12", "--change-cutoff-to-inf-under-threshold", thr)
    _compare(o, xg, lg)
    o.close()


def test_sub_configurations(tmp_path):
    xyb, d = _setup(tmp_path)
    sub = tmp_path / "sub"
    sub.mkdir()
    o = _oracle(xyb)
    o.run(step_count_limit=6, sub_configs=True, sub_delay=5, sub_delay_avalanche=1)
    times, xs = o.subconfigs()
    # a counter that restarts at 0 never exceeds the delay: one file per step
    assert len(times) == 6
    xg, lg, _ = _cli(tmp_path, d, "--step-count-limit", "6", "--save-sub-configurations", str(sub))
    _compare(o, xg, lg)
    names = sorted(os.listdir(sub))
    assert len(names) == len(times)
    got = {}
    for nm in names:
        assert nm.endswith(".dconf")
        got[float(nm[:-6])] = configs.read_dconf(str(sub / nm))
    for t, x in zip(times, xs):
        # the file name is the time printed with 16 significant digits
        key = min(got, key=lambda u: abs(u - t))
        assert abs(key - t) <= 1e-12 * t
        assert ("%.16g" % t) == ("%.16g" % key) or abs(key - t) <= 1e-13 * t
        gx, gy, gb = got[key]
        assert np.max(np.abs(gx - x)) < 1e-9 and np.array_equal(gy, xyb[1]) and np.array_equal(gb, xyb[2])
    o.close()


def test_spring_protocol(tmp_path):
    # sigma = k (rate t - strain), clamped at 0, with the stage strains
    # (spring_protocol.cpp:33-59, simulation.cpp:151-184)
    xyb, d = _setup(tmp_path)
    o = _oracle(xyb, stress="spring", p1=200.0, p2=0.5)
    o.run(step_count_limit=10)
    xg, lg, _ = _cli(tmp_path, d, "--step-count-limit", "10", "--fixed-rate-external-stress", "200",
                     "--spring-constant", "0.5")
    _compare(o, xg, lg)
    assert np.all(np.diff(lg[1:, 7]) > 0)
    o.close()


def test_forced_rejection_from_large_first_step(tmp_path):
    # h0 = 1e-3 is far above what precision 1e-6 allows: the first attempt is
    # rejected and the controller shrinks h (simulation.cpp:638-649); the
    # rejected big step stays finite and within 21 cells (checked on the
    # oracle before the device runs it)
    xyb, d = _setup(tmp_path)
    o = _oracle(xyb, h0=1e-3)
    o.step()
    assert o.failed_steps == 1 and np.all(np.isfinite(o.big_x())) and np.abs(o.big_x()).max() < 25
    o.run(step_count_limit=5)
    xg, lg, _ = _cli(tmp_path, d, "--step-count-limit", "5", "--initial-stepsize", "1e-3")
    _compare(o, xg, lg)
    assert lg[-1, 2] >= 1
    o.close()


def test_coincident_pair_tolerance():
    # two dislocations at the same place: their tolerances become EPS = 1e-12
    # (precision_handler.cpp:77-81) and their pair term is 0/0 (NaN speeds);
    # every other speed stays finite and equal to the oracle's
    n = 64
    x, y, b = configs.dislocations(n, 5)
    x[7], y[7] = x[3], y[3]
    for mode in (S.FIELD_EXACT, S.FIELD_ORDERED, S.FIELD_FAST):
        with S.Context(n, precision=1e-5, field_mode=mode) as c:
            c.upload_state(0, x, y, b)
            c.reset_tolerance()
            v = c.speeds(0, S.V_INIT, 0.0)
            tol = c.tolerance()
        vr, tr = O.speeds(x, y, b, prec=1e-5)
        assert tr[3] == 1e-12 and tr[7] == 1e-12
        assert np.array_equal(tol.view(np.int64), tr.view(np.int64))
        assert np.array_equal(np.isnan(v), np.isnan(vr)) and np.isnan(v[3]) and np.isnan(v[7])
        fin = ~np.isnan(vr)
        assert np.max(np.abs(v[fin] - vr[fin]) / np.abs(vr[fin])) < 1e-11
