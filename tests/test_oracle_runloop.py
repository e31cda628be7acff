"""Pins of the oracle's run loop (orc_sim_run; reference simulation.cpp:
417-433 and the accept branch 553-633) on the CPU, against an independent
record: the log of an unconstrained oracle run of the same configuration.
Each limit is placed halfway between two consecutive logged values, and the
loop's documented semantics -- the conditions are tested before every
attempt, the speed and stress limits and the avalanche counter act on the
accepted step's logged mean |v| and stress, the cutoff switch from the next
step on, post-relax adds whole successful steps -- fix the step at which the
constrained run must end and what it must have logged.  An off-by-one in any
condition, a >= for a >, a counter advanced on the wrong edge or a switch
applied a step early fails one of these.  The device runs are compared with
the same oracle in test_gpu_runloop.py."""
import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import configs

N, SEED, M = 96, 71, 1.0      # m = 1 < sqrt(2N)/12: a finite cutoff


def _xyb():
    return configs.dislocations(N, SEED)


def _oracle(stress="zero", p1=0.0):
    x, y, b = _xyb()
    return O.OracleSim(x, y, b, cutoff_multiplier=M, prec=1e-6, h0=1e-6, stress=stress, p1=p1,
                       calculate_strain=True)


def _free(steps, **kw):
    o = _oracle(**kw)
    o.run(step_count_limit=steps)
    log, x = o.log(), o.x()
    o.close()
    return log, x


def _mid(log, row, col):
    return 0.5 * (log[row, col] + log[row + 1, col])


def _same_prefix(log, free):
    # the constrained run is the free run cut short: identical logged rows
    k = len(log)
    cols = [0, 1, 2, 3, 4, 5, 7, 9, 10, 11]
    assert np.array_equal(log[:, cols], free[:k, cols], equal_nan=True)


def test_step_count_limit_is_exact_and_changes_nothing():
    free, _ = _free(12)
    o = _oracle()
    o.run(step_count_limit=7)
    assert o.succesful_steps == 7
    _same_prefix(o.log(), free)
    # seven more from where it stopped reach the free run's state bit for bit
    o.run(step_count_limit=12)
    _, x12 = _free(12)
    assert np.array_equal(o.x(), x12)
    o.close()


def test_time_limit_stops_at_first_time_beyond():
    free, _ = _free(12)
    thr = _mid(free, 8, 0)
    o = _oracle()
    o.run(time_limit=thr)
    log = o.log()
    # the loop tests simTime < limit before each attempt: the last logged
    # time is the first one at or beyond the limit
    assert log[-1, 0] >= thr > log[-2, 0] and o.succesful_steps == free[8 + 1, 1]
    _same_prefix(log, free)
    o.close()


def test_strain_limit_stops_at_first_strain_beyond():
    free, _ = _free(12, stress="constant", p1=40.0)
    thr = _mid(free, 7, 9)
    o = _oracle(stress="constant", p1=40.0)
    o.run(strain_limit=thr)
    log = o.log()
    assert log[-1, 9] >= thr > log[-2, 9]
    _same_prefix(log, free)
    o.close()


def test_speed_limit_finishes_on_the_slow_step_then_post_relax():
    free, _ = _free(16)
    thr = _mid(free, 10, 4)
    slow = np.nonzero(free[:, 4] < thr)[0][0]          # first accepted step below the limit
    o = _oracle()
    o.run(speed_limit=thr, post_relax=3)
    log = o.log()
    assert o.run_state()["finish"] == 1.0 and o.run_state()["remaining_final_steps"] == 0.0
    # finish is raised on the slow step's acceptance; post-relax adds three
    # successful steps after it
    assert o.succesful_steps == free[slow, 1] + 3
    _same_prefix(log, free)
    o.close()


def test_stress_limit_finishes_on_the_step_above():
    free, _ = _free(12, stress="fixed-rate", p1=200.0)
    thr = _mid(free, 9, 7)
    above = np.nonzero(free[:, 7] > thr)[0][0]
    o = _oracle(stress="fixed-rate", p1=200.0)
    o.run(stress_limit=thr)
    assert o.run_state()["finish"] == 1.0 and o.succesful_steps == free[above, 1]
    _same_prefix(o.log(), free)
    o.close()


@pytest.mark.parametrize("row", [3, 8])
def test_avalanche_count_is_falling_edges_of_mean_speed(row):
    # the relaxing configuration's mean |v| falls through the threshold once
    free, _ = _free(16)
    v = free[:, 4]
    thr = _mid(free, row, 4)
    limit = 1
    # an avalanche is under way while the accepted steps' mean |v| is above
    # the threshold and counts once when it drops below it again
    count, inside, stop = 0, False, None
    for k, s in enumerate(v):
        if inside and s < thr:
            count += 1
            inside = False
            if count == limit:
                stop = k
                break
        elif s > thr:
            inside = True
    assert stop == row + 1
    o = _oracle()
    o.run(avalanche_limit=limit, avalanche_threshold=thr)
    st = o.run_state()
    assert st["avalanche_count"] == limit and st["in_avalanche"] == 0.0
    assert o.succesful_steps == free[stop, 1]
    _same_prefix(o.log(), free)
    o.close()


def test_cutoff_switch_from_the_step_after_the_slow_one():
    free, _ = _free(12)
    thr = _mid(free, 6, 4)
    slow = np.nonzero(free[:, 4] < thr)[0][0]
    o = _oracle()
    o.run(step_count_limit=12, cutoff_inf_threshold=thr)
    log = o.log()
    rc = 1.0 / np.sqrt(N) * M
    # the slow step still logs the old cutoff; the multiplier is 1e20 after it
    assert np.all(log[: slow + 1, 5] == rc)
    assert np.all(log[slow + 1:, 5] == 1.0 / np.sqrt(N) * 1e20)
    assert o.run_state()["cutoff_multiplier"] == 1e20
    assert np.array_equal(log[: slow + 1, [0, 4]], free[: slow + 1, [0, 4]])
    o.close()


def test_sub_configurations_on_every_accepted_step():
    # the write condition delay >= counter holds at counter 0, and the counter
    # is reset by every write, so each accepted step writes its positions,
    # named by its time
    o = _oracle()
    o.run(step_count_limit=6, sub_configs=True, sub_delay=5, sub_delay_avalanche=1)
    t, xs = o.subconfigs()
    log = o.log()
    # (the log's first row is the initial configuration's, at t = 0)
    assert len(t) == 6 and np.array_equal(t, log[1:, 0]) and log[0, 0] == 0.0
    assert np.array_equal(xs[-1], o.x())
    o.close()
