"""Python surface of the C++ simulator API (libsdddstCore.so): the role of the
reference's PySdddstCore module (bindings.cpp:140-206) — build a system,
step it, run it, read the time, the positions and the 12-column log."""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._capi import SdddstError, load_library

_HERE = os.path.dirname(os.path.abspath(__file__))
CORE_PATH = os.path.join(_HERE, "libsdddstCore.so")

PROTOCOLS = {"zero": 0, "fixed-rate": 1, "spring": 2, "constant": 3}
SOLVERS = {"auto": 0, "umfpack": 1, "dense-lu": 2, "fgmres": 3}
FIELD_MODES = {"fast": 0, "exact": 1, "ordered": 2}
PRECONDITIONERS = {"jacobi": 0, "pair": 1, "learned": 2, "strong": 3}

_dp = ctypes.POINTER(ctypes.c_double)


class SimOptions(ctypes.Structure):
    _fields_ = [
        ("cutoff_multiplier", ctypes.c_double),
        ("precision", ctypes.c_double),
        ("initial_step", ctypes.c_double),
        ("max_step", ctypes.c_double),
        ("protocol", ctypes.c_int),
        ("stress_param1", ctypes.c_double),
        ("stress_param2", ctypes.c_double),
        ("calculate_strain", ctypes.c_int),
        ("order_parameter", ctypes.c_int),
        ("linear_solver", ctypes.c_int),
        ("solver_tol", ctypes.c_double),
        ("solver_max_iter", ctypes.c_int),
        ("solver_restart", ctypes.c_int),
        ("field_mode", ctypes.c_int),
        ("preconditioner", ctypes.c_int),
        ("A", ctypes.c_double),
        ("KASQR", ctypes.c_double),
        ("log_path", ctypes.c_char_p),
        ("spai_weights", ctypes.c_char_p),
    ]


class SimState(ctypes.Structure):
    _fields_ = [
        ("sim_time", ctypes.c_double),
        ("step_size", ctypes.c_double),
        ("strain", ctypes.c_double),
        ("energy", ctypes.c_double),
        ("max_error_ratio_sqr", ctypes.c_double),
        ("sum_avg_speed", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("succesful_steps", ctypes.c_longlong),
        ("failed_steps", ctypes.c_longlong),
        ("last_step_succeeded", ctypes.c_int),
        ("last_step_wall_s", ctypes.c_double),
        ("solver_last_iters", ctypes.c_int),
        ("solver_fallbacks", ctypes.c_int),
        ("solver_last_rel_residual", ctypes.c_double),
        ("solver_total_iters", ctypes.c_longlong),
        ("solver_total_solves", ctypes.c_longlong),
        ("solver_precond_fallbacks", ctypes.c_int),
        ("aborted_attempts", ctypes.c_longlong),
        ("solver_strong_retries", ctypes.c_int),
    ]


_core = None


def load_core():
    global _core
    if _core is not None:
        return _core
    load_library()  # the C ABI library first (dependency)
    if not os.path.exists(CORE_PATH):
        raise SdddstError(f"{CORE_PATH} is missing: run `make -C sdddst_amd`")
    lib = ctypes.CDLL(CORE_PATH)
    vp = ctypes.c_void_p
    sig = {
        "sdddst_sim_default_options": (None, [ctypes.POINTER(SimOptions)]),
        "sdddst_sim_create": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp,
                                             ctypes.POINTER(SimOptions), ctypes.POINTER(vp)]),
        "sdddst_sim_create_from_files": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p,
                                                        ctypes.POINTER(SimOptions), ctypes.POINTER(vp)]),
        "sdddst_sim_destroy": (ctypes.c_int, [vp]),
        "sdddst_sim_step": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
        "sdddst_sim_run": (ctypes.c_int, [vp, ctypes.c_int]),
        "sdddst_sim_set_step_size": (ctypes.c_int, [vp, ctypes.c_double]),
        "sdddst_sim_get_state": (ctypes.c_int, [vp, ctypes.POINTER(SimState)]),
        "sdddst_sim_positions": (ctypes.c_int, [vp, _dp, ctypes.c_int]),
        "sdddst_sim_flush_log": (ctypes.c_int, [vp]),
        "sdddst_sim_write_dconf": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sdddst_sim_solver_name": (ctypes.c_char_p, [vp]),
        "sdddst_core_last_error": (ctypes.c_char_p, []),
        "sdddst_sim_device_context": (ctypes.c_void_p, [vp]),
        "sdddst_core_field_eval": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _core = lib
    return lib


def analytic_field(which: int, dx, dy) -> np.ndarray:
    """AnalyticField's Field virtuals through the C++ class (which = 0 xy,
    1 xy_diff_x, 2 xy_diff_x2), per point."""
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    out = np.empty_like(dx)
    lib = load_core()
    if lib.sdddst_core_field_eval(which, len(dx), dx.ctypes.data_as(_dp), dy.ctypes.data_as(_dp),
                                  out.ctypes.data_as(_dp)) != 0:
        raise SdddstError(lib.sdddst_core_last_error().decode())
    return out


def default_options() -> SimOptions:
    o = SimOptions()
    load_core().sdddst_sim_default_options(ctypes.byref(o))
    return o


@dataclass
class StepState:
    sim_time: float
    step_size: float
    strain: float
    energy: float
    max_error_ratio_sqr: float
    sum_avg_speed: float
    cutoff: float
    succesful_steps: int
    failed_steps: int
    last_step_succeeded: bool
    last_step_wall_s: float
    solver_last_iters: int
    solver_fallbacks: int
    solver_last_rel_residual: float
    solver_total_iters: int
    solver_total_solves: int
    solver_precond_fallbacks: int
    aborted_attempts: int = 0
    solver_strong_retries: int = 0


class Simulation:
    """sdddstCore::Simulation over a SimulationData built from arrays."""

    def __init__(self, x, y, b, *, cutoff_multiplier=1.0, precision=1e-6, initial_step=1e-6, max_step=None,
                 stress="zero", stress_value=0.0, spring_constant=1.0, calculate_strain=True,
                 order_parameter=False, solver="auto", solver_tol=1e-10, solver_max_iter=300, solver_restart=40,
                 field_mode="fast", preconditioner="pair", px=None, py=None, A=1e-4 * 16.0, KASQR=1.65 * 1.65 * 1e6 / 256.0,
                 log_path: Optional[str] = None, spai_weights: Optional[str] = None):
        self.lib = load_core()
        o = default_options()
        o.cutoff_multiplier = cutoff_multiplier
        o.precision = precision
        o.initial_step = initial_step
        o.max_step = -1.0 if max_step is None else max_step
        o.protocol = PROTOCOLS[stress]
        o.stress_param1 = stress_value
        o.stress_param2 = spring_constant
        o.calculate_strain = int(calculate_strain)
        o.order_parameter = int(order_parameter)
        o.linear_solver = SOLVERS[solver]
        o.solver_tol = solver_tol
        o.solver_max_iter = solver_max_iter
        o.solver_restart = solver_restart
        o.field_mode = FIELD_MODES[field_mode]
        o.preconditioner = PRECONDITIONERS[preconditioner]
        o.A = A
        o.KASQR = KASQR
        self._log = log_path.encode() if log_path else None
        o.log_path = self._log
        self._weights = spai_weights.encode() if spai_weights else None
        o.spai_weights = self._weights
        self.n = len(x)
        xs, ys, bs = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b))
        pc = 0 if px is None else len(px)
        pxs = None if px is None else np.ascontiguousarray(px, dtype=np.float64)
        pys = None if py is None else np.ascontiguousarray(py, dtype=np.float64)
        h = ctypes.c_void_p()
        rc = self.lib.sdddst_sim_create(self.n, xs.ctypes.data_as(_dp), ys.ctypes.data_as(_dp),
                                        bs.ctypes.data_as(_dp), pc,
                                        None if pxs is None else pxs.ctypes.data_as(_dp),
                                        None if pys is None else pys.ctypes.data_as(_dp), ctypes.byref(o),
                                        ctypes.byref(h))
        self._check(rc)
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise SdddstError(f"status {rc}: {self.lib.sdddst_core_last_error().decode(errors='replace')}")

    def step(self) -> bool:
        ok = ctypes.c_int()
        self._check(self.lib.sdddst_sim_step(self.h, ctypes.byref(ok)))
        return bool(ok.value)

    def run(self, successful_steps: int):
        self._check(self.lib.sdddst_sim_run(self.h, successful_steps))

    def set_step_size(self, h: float):
        self._check(self.lib.sdddst_sim_set_step_size(self.h, h))

    def state(self) -> StepState:
        s = SimState()
        self._check(self.lib.sdddst_sim_get_state(self.h, ctypes.byref(s)))
        return StepState(s.sim_time, s.step_size, s.strain, s.energy, s.max_error_ratio_sqr, s.sum_avg_speed,
                         s.cutoff, s.succesful_steps, s.failed_steps, bool(s.last_step_succeeded),
                         s.last_step_wall_s, s.solver_last_iters, s.solver_fallbacks, s.solver_last_rel_residual,
                         s.solver_total_iters, s.solver_total_solves, s.solver_precond_fallbacks,
                         s.aborted_attempts, s.solver_strong_retries)

    def get_time(self) -> float:
        return self.state().sim_time

    def positions(self) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self.lib.sdddst_sim_positions(self.h, out.ctypes.data_as(_dp), self.n))
        return out

    def flush_log(self):
        self._check(self.lib.sdddst_sim_flush_log(self.h))

    def write_dconf(self, path: str):
        self._check(self.lib.sdddst_sim_write_dconf(self.h, path.encode()))

    @property
    def device_context(self):
        """Handle of the sdddst_ctx (for _capi.timing / kernel_stats)."""
        return ctypes.c_void_p(self.lib.sdddst_sim_device_context(self.h))

    @property
    def solver_name(self) -> str:
        return self.lib.sdddst_sim_solver_name(self.h).decode()

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdddst_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_log(path: str) -> np.ndarray:
    """The 12-column log as floats ('-' columns become NaN)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) != 12:
                continue
            rows.append([float("nan") if p == "-" else float(p) for p in parts])
    return np.array(rows).reshape(-1, 12)
