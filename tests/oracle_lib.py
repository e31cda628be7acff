"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when it was
built, to the reference's own translation units (oracle/_ref).  Test
infrastructure: used by the tests, the smoke check and the benchmark's
CPU-baseline leg only.  Builds the oracle with make when it is missing."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# SDDDST_ORACLE_SO points the tests at another build of the oracle (a copy
# with a planted mistake, to see which pins catch it)
ORACLE_SO = os.environ.get("SDDDST_ORACLE_SO", os.path.join(ORACLE_DIR, "liboracle.so"))
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libsdddst_ref.so")
MATHCHECK_SO = os.path.join(ROOT, "tests", "native", "libmathcheck.so")

D = ctypes.c_double
DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)
L = ctypes.c_long
VP = ctypes.c_void_p


def dp(a):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(DP)


def ip(a):
    return None if a is None else a.ctypes.data_as(IP)


def _build(target_dir, target):
    subprocess.run(["make", "-s", "-C", target_dir, target], check=True, stdout=subprocess.DEVNULL)


_oracle = None


def oracle():
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(ORACLE_SO):
        _build(ORACLE_DIR, "oracle")
    lib = ctypes.CDLL(ORACLE_SO)
    sig = {
        "orc_normalize": (D, [D]),
        "orc_xy": (D, [D, D]),
        "orc_xy_diff_x": (D, [D, D]),
        "orc_xy_diff_x2": (D, [D, D]),
        "orc_new_step_size": (D, [D, D]),
        "orc_ph_update_tolerance": (None, [DP, D, D, ctypes.c_uint]),
        "orc_max_error_ratio": (D, [ctypes.c_int, DP, DP, DP, IP]),
        "orc_speeds": (None, [ctypes.c_int, DP, DP, DP, ctypes.c_int, DP, DP, D, D, D, DP, DP, D]),
        "orc_speeds_rows_sample": (D, [ctypes.c_int, DP, DP, DP, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_long)]),
        "orc_cutoff": (None, [ctypes.c_int, D, DP, DP, DP]),
        "orc_jacobian": (L, [ctypes.c_int, DP, DP, DP, ctypes.c_int, DP, DP, D, D, D, D, L, IP, IP, DP, DP, IP,
                             DP]),
        "orc_dense_solve_csc": (ctypes.c_int, [ctypes.c_int, IP, IP, DP, DP, DP]),
        "orc_sim_create": (VP, [ctypes.c_int, DP, DP, DP, ctypes.c_int, DP, DP, D, D, D, D, D, ctypes.c_int, D, D,
                                ctypes.c_int, L]),
        "orc_sim_destroy": (None, [VP]),
        "orc_sim_step": (ctypes.c_int, [VP]),
        "orc_sim_set_step_size": (None, [VP, D]),
        "orc_sim_set_order_parameter": (None, [VP, ctypes.c_int]),
        "orc_sim_set_max_step_size": (None, [VP, D]),
        "orc_sim_get_step_size": (D, [VP]),
        "orc_sim_get_time": (D, [VP]),
        "orc_sim_succesful_steps": (L, [VP]),
        "orc_sim_failed_steps": (L, [VP]),
        "orc_sim_energy": (D, [VP]),
        "orc_sim_strain": (D, [VP]),
        "orc_sim_max_error_ratio_last": (D, [VP]),
        "orc_sim_get_x": (None, [VP, DP]),
        "orc_sim_log_rows": (L, [VP]),
        "orc_sim_get_log": (None, [VP, DP]),
        "orc_sim_get_speeds": (None, [VP, DP, DP]),
        "orc_sim_get_big_x": (None, [VP, DP]),
        "orc_sim_get_fs_x": (None, [VP, DP]),
        "orc_sim_run": (L, [VP, ctypes.c_void_p]),
        "orc_sim_run_state": (None, [VP, DP]),
        "orc_sim_subconfigs": (L, [VP, DP, DP, L]),
        "orc_sim_residual": (None, [VP, D, DP, DP, DP, DP, DP]),
        "orc_sim_stage": (ctypes.c_int, [VP, D, DP, DP]),
        "orc_libm_batch": (None, [ctypes.c_int, L, DP, DP]),
        "orc_xy_batch": (None, [L, DP, DP, DP, DP]),
        "orc_speeds_abs": (None, [ctypes.c_int, DP, DP, DP, DP]),
        "orc_pd_diag_term": (D, [D, D, D, D, D, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _oracle = lib
    return lib


_ref = None


def reference():
    """The reference's own AnalyticField / PrecisionHandler build, or None."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    sig = {
        "ref_normalize": (D, [D]),
        "ref_xy": (D, [D, D]),
        "ref_xy_diff_x": (D, [D, D]),
        "ref_xy_diff_x2": (D, [D, D]),
        "ref_xy_batch": (None, [L, DP, DP, DP, DP]),
        "ref_speeds": (None, [ctypes.c_int, DP, DP, DP, D, DP]),
        "ref_precision_sequence": (D, [ctypes.c_uint, D, ctypes.c_int, DP, ctypes.POINTER(ctypes.c_uint),
                                       ctypes.c_int, DP, ctypes.POINTER(ctypes.c_uint), D, DP]),
        "ref_protocol_stress": (D, [ctypes.c_int, D, D, D, D, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _ref = lib
    return lib


_mc = None


def mathcheck():
    """Host build of the device math (tests/native)."""
    global _mc
    if _mc is not None:
        return _mc
    # make rebuilds it when the device headers it compiles have changed
    _build(os.path.join(ROOT, "tests", "native"), "libmathcheck.so")
    lib = ctypes.CDLL(MATHCHECK_SO)
    lib.mc_batch1.argtypes = [ctypes.c_int, L, DP, DP]
    lib.mc_batch2.argtypes = [ctypes.c_int, L, DP, DP, DP]
    lib.mc_batch_pair_far.argtypes = [L, DP, DP, DP, DP, DP]
    lib.mc_batch_pair_gap.argtypes = [L, DP, DP, DP, DP, DP]
    lib.mc_batch_pair_mixed.argtypes = [L, DP, DP, DP, DP, DP, IP]
    lib.mc_tile_kind.argtypes = [DP, DP, DP, DP]
    lib.mc_tile_kind.restype = ctypes.c_int
    lib.mc_normalize.restype = D
    lib.mc_normalize.argtypes = [D]
    lib.mc_cos_2pi.restype = D
    lib.mc_cos_2pi.argtypes = [D]
    lib.mc_powi.restype = D
    lib.mc_powi.argtypes = [D, ctypes.c_int]
    _mc = lib
    return lib


# -- convenience wrappers --------------------------------------------------

def libm_batch(fn: int, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    oracle().orc_libm_batch(fn, len(x), dp(x), dp(out))
    return out


def xy_batch(dx, dy):
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    a, b = np.empty_like(dx), np.empty_like(dx)
    oracle().orc_xy_batch(len(dx), dp(dx), dp(dy), dp(a), dp(b))
    return a, b


def speeds(x, y, b, sigma=0.0, px=None, py=None, A=0.0, K=0.0, tol=None, prec=1e-6):
    """orc_speeds: returns (v, tol) with tol updated from its input (or from
    prec^2 when tol is None)."""
    n = len(x)
    v = np.empty(n)
    t = np.full(n, prec * prec) if tol is None else np.array(tol, dtype=np.float64)
    pc = 0 if px is None else len(px)
    oracle().orc_speeds(n, dp(x), dp(y), dp(b), pc, dp(px), dp(py), A, K, sigma, dp(v), dp(t), prec * prec)
    return v, t


def speeds_abs(x, y, b):
    """S_i = sum_j |t_ij| per target."""
    n = len(x)
    out = np.empty(n)
    oracle().orc_speeds_abs(n, dp(x), dp(y), dp(b), dp(out))
    return out


def speed_parity(v, vr, S, n):
    """Per-component speed parity: within 1e-12 relative, or, where the pair
    sum cancels, within the rounding scale of a reordered n-term sum plus the
    per-pair absolute error floor: 2^-52 * sqrt(n) * (S_i + n)."""
    dv = np.abs(np.asarray(v) - np.asarray(vr))
    rel_ok = dv <= 1e-12 * np.abs(vr)
    cond_ok = dv <= 2.0 ** -52 * np.sqrt(n) * (S + n)
    return rel_ok, rel_ok | cond_ok


def jacobian(x, y, b, cutoff_multiplier, h, cap=None, px=None, py=None, A=0.0, K=0.0):
    n = len(x)
    pc = 0 if px is None else len(px)
    cap = cap or min(n * n, 400 * n + 4096)
    while True:
        Ap = np.zeros(n + 1, dtype=np.int32)
        Ai = np.zeros(cap, dtype=np.int32)
        Ax = np.zeros(cap)
        dv = np.zeros(n)
        idx = np.zeros(n, dtype=np.int32)
        raw = np.zeros(cap)
        nnz = oracle().orc_jacobian(n, dp(x), dp(y), dp(b), pc, dp(px), dp(py), A, K, cutoff_multiplier, h, cap,
                                    ip(Ap), ip(Ai), dp(Ax), dp(dv), ip(idx), dp(raw))
        if nnz >= 0:
            return Ap, Ai[:nnz].copy(), Ax[:nnz].copy(), dv, idx, raw[:nnz].copy()
        cap = min(n * n, cap * 2)


def cutoff(n, m):
    a, b, c = D(), D(), D()
    oracle().orc_cutoff(n, m, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


class RunOpts(ctypes.Structure):
    """orc_run_opts: the run-loop settings (simulation.cpp:417-433, 553-633)."""
    _fields_ = [("isTimeLimit", ctypes.c_int), ("timeLimit", D),
                ("isStrainIncreaseLimit", ctypes.c_int), ("strainIncreaseLimit", D),
                ("isStepCountLimit", ctypes.c_int), ("stepCountLimit", L),
                ("countAvalanches", ctypes.c_int), ("avalancheTriggerLimit", L), ("avalancheSpeedThreshold", D),
                ("isSpeedLimit", ctypes.c_int), ("speedLimit", D),
                ("isStressLimit", ctypes.c_int), ("stressLimit", D),
                ("isSpeedThresholdForCutoffChange", ctypes.c_int), ("speedThresholdForCutoffChange", D),
                ("isSaveSubConfigs", ctypes.c_int), ("subConfigDelay", L), ("subConfigDelayDuringAvalanche", L),
                ("remainingFinalSteps", L)]


class OracleSim:
    """The oracle's full stepping driver (dense LU stands in for UMFPACK)."""

    PROTOCOLS = {"zero": 0, "fixed-rate": 1, "spring": 2, "constant": 3}

    def __init__(self, x, y, b, cutoff_multiplier=1.0, prec=1e-6, h0=1e-6, stress="zero", p1=0.0, p2=0.0,
                 calculate_strain=True, order_parameter=False, px=None, py=None, A=0.0, K=0.0, cap=0):
        self.lib = oracle()
        self.n = len(x)
        pc = 0 if px is None else len(px)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b)]
        self.h = self.lib.orc_sim_create(self.n, dp(self._keep[0]), dp(self._keep[1]), dp(self._keep[2]), pc,
                                         dp(px), dp(py), A, K, cutoff_multiplier, prec, h0,
                                         self.PROTOCOLS[stress], p1, p2, int(calculate_strain), cap)
        if order_parameter:
            self.lib.orc_sim_set_order_parameter(self.h, 1)

    def step(self) -> int:
        rc = self.lib.orc_sim_step(self.h)
        if rc < 0:
            raise RuntimeError("oracle step failed")
        return rc

    def set_step_size(self, h):
        self.lib.orc_sim_set_step_size(self.h, h)

    @property
    def step_size(self):
        return self.lib.orc_sim_get_step_size(self.h)

    @property
    def time(self):
        return self.lib.orc_sim_get_time(self.h)

    @property
    def succesful_steps(self):
        return self.lib.orc_sim_succesful_steps(self.h)

    @property
    def failed_steps(self):
        return self.lib.orc_sim_failed_steps(self.h)

    def x(self):
        out = np.empty(self.n)
        self.lib.orc_sim_get_x(self.h, dp(out))
        return out

    def big_x(self):
        out = np.empty(self.n)
        self.lib.orc_sim_get_big_x(self.h, dp(out))
        return out

    def run(self, time_limit=None, strain_limit=None, step_count_limit=None, avalanche_limit=None,
            avalanche_threshold=1e-3, speed_limit=None, stress_limit=None, cutoff_inf_threshold=None,
            sub_configs=False, sub_delay=5, sub_delay_avalanche=1, post_relax=0):
        """Simulation::run with the command line's settings; returns the
        number of attempts."""
        o = RunOpts()
        for value, value_field, flag_field in ((time_limit, "timeLimit", "isTimeLimit"),
                                (strain_limit, "strainIncreaseLimit", "isStrainIncreaseLimit"),
                                (step_count_limit, "stepCountLimit", "isStepCountLimit"),
                                (avalanche_limit, "avalancheTriggerLimit", "countAvalanches"),
                                (speed_limit, "speedLimit", "isSpeedLimit"),
                                (stress_limit, "stressLimit", "isStressLimit"),
                                (cutoff_inf_threshold, "speedThresholdForCutoffChange",
                                 "isSpeedThresholdForCutoffChange")):
            if value is not None:
                setattr(o, flag_field, 1)
                setattr(o, value_field, value)
        o.avalancheSpeedThreshold = avalanche_threshold
        o.isSaveSubConfigs = int(sub_configs)
        o.subConfigDelay = sub_delay
        o.subConfigDelayDuringAvalanche = sub_delay_avalanche
        o.remainingFinalSteps = post_relax
        r = self.lib.orc_sim_run(self.h, ctypes.addressof(o))
        if r < 0:
            raise RuntimeError("oracle run failed")
        return r

    def run_state(self):
        out = np.empty(7)
        self.lib.orc_sim_run_state(self.h, dp(out))
        keys = ("avalanche_count", "in_avalanche", "finish", "remaining_final_steps", "cutoff", "cutoff_multiplier",
                "subconfig_distance_counter")
        return dict(zip(keys, out.tolist()))

    def subconfigs(self):
        k = self.lib.orc_sim_subconfigs(self.h, None, None, 0)
        t, xs = np.empty(k), np.empty((k, self.n))
        self.lib.orc_sim_subconfigs(self.h, dp(t), dp(xs), k)
        return t, xs

    def fs_x(self):
        out = np.empty(self.n)
        self.lib.orc_sim_get_fs_x(self.h, dp(out))
        return out

    def residual(self, h, newx, oldx, dvec, v_old):
        """One calculateG of the stage-I kind with dVec and the old speeds given."""
        g = np.empty(self.n)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (newx, oldx, dvec, v_old)]
        self.lib.orc_sim_residual(self.h, h, *(dp(a) for a in args), dp(g))
        return g

    def stage(self, h):
        """One integrate of the stage-I kind from the current positions:
        returns (end positions, d vector)."""
        newx, dv = np.empty(self.n), np.empty(self.n)
        if self.lib.orc_sim_stage(self.h, h, dp(newx), dp(dv)) != 0:
            raise RuntimeError("oracle integrate failed")
        return newx, dv

    def log(self):
        rows = self.lib.orc_sim_log_rows(self.h)
        out = np.empty((rows, 12))
        if rows:
            self.lib.orc_sim_get_log(self.h, dp(out))
        return out

    def close(self):
        if self.h:
            self.lib.orc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
