"""Batched ensembles of independent realizations on one device (ctypes
binding of the sdddst_ensemble_* entry points of include/sdddst_amd.h).

Every realization follows sdddstCore::Simulation::step under a constant (or
zero) external stress; the realizations advance in lock step, one attempt
of all of them per replay of a captured HIP graph.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._capi import FIELD_EXACT, FIELD_FAST, Params, SdddstError, load_library

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

_SIGS = {
    "sdddst_ensemble_create": [ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.POINTER(Params), ctypes.c_double,
                               ctypes.c_double, ctypes.POINTER(_vp)],
    "sdddst_ensemble_destroy": [_vp],
    "sdddst_ensemble_set_tiles": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "sdddst_ensemble_autotune": [_vp, ctypes.c_char_p, ctypes.c_int],
    "sdddst_ensemble_run": [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _i64p],
    "sdddst_ensemble_state": [_vp, _dp, _dp, _i64p, _i64p],
    "sdddst_ensemble_positions": [_vp, _dp],
    "sdddst_ensemble_stats": [_vp, _i64p, _i64p, _i64p, _i64p, _i64p, _ip, _ip, _ip],
}


def _lib():
    lib = load_library()
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    return lib


@dataclass
class EnsembleStats:
    attempts: int
    solver_iterations: int
    solves: int
    lu_fallbacks: int
    regrows: int
    speeds_block: int
    solver_block: int
    spmv_lanes: int


class Ensemble:
    """num realizations of sizes n (x, y, b concatenated in that order)."""

    def __init__(self, sizes, x, y, b, *, cutoff_multiplier=1.0, precision=1e-6, field_mode="fast", sigma=0.0,
                 initial_step=1e-6, iteration_count=2):
        self.lib = _lib()
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self.nt = int(self.sizes.sum())
        xs, ys, bs = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b))
        if not (len(xs) == len(ys) == len(bs) == self.nt):
            raise ValueError("x, y, b must hold sum(sizes) values")
        mode = {"fast": FIELD_FAST, "exact": FIELD_EXACT}[field_mode]
        p = Params(cutoff_multiplier, precision, 0.0, 0.0, iteration_count, mode)
        h = _vp()
        self._check(self.lib.sdddst_ensemble_create(len(self.sizes), self.sizes.ctypes.data_as(_ip),
                                                    xs.ctypes.data_as(_dp), ys.ctypes.data_as(_dp),
                                                    bs.ctypes.data_as(_dp), ctypes.byref(p), sigma, initial_step,
                                                    ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise SdddstError(f"status {rc}: {self.lib.sdddst_last_error().decode(errors='replace')}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdddst_ensemble_destroy(self.h)
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

    def set_tiles(self, speeds_block: int, solver_block: int, spmv_lanes: int = 8):
        """speeds_block 0: symmetric tiles; else one lane per target."""
        self._check(self.lib.sdddst_ensemble_set_tiles(self.h, speeds_block, solver_block, spmv_lanes))

    def autotune(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        self._check(self.lib.sdddst_ensemble_autotune(self.h, buf, len(buf)))
        return buf.value.decode()

    def run(self, target_steps: int, max_attempts: int = 1 << 40, use_graph: bool = True) -> int:
        a = ctypes.c_int64()
        self._check(self.lib.sdddst_ensemble_run(self.h, target_steps, max_attempts, int(use_graph), ctypes.byref(a)))
        return a.value

    def state(self):
        """(time, step size, successful steps, failed steps) per realization."""
        k = len(self.sizes)
        t, h = np.empty(k), np.empty(k)
        s, f = np.empty(k, np.int64), np.empty(k, np.int64)
        self._check(self.lib.sdddst_ensemble_state(self.h, t.ctypes.data_as(_dp), h.ctypes.data_as(_dp),
                                                   s.ctypes.data_as(_i64p), f.ctypes.data_as(_i64p)))
        return t, h, s, f

    def positions(self) -> np.ndarray:
        x = np.empty(self.nt)
        self._check(self.lib.sdddst_ensemble_positions(self.h, x.ctypes.data_as(_dp)))
        return x

    def split(self, arr):
        return np.split(np.asarray(arr), np.cumsum(self.sizes)[:-1])

    def stats(self) -> EnsembleStats:
        v = [ctypes.c_int64() for _ in range(5)]
        sb, vb, lb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.sdddst_ensemble_stats(self.h, *(ctypes.byref(q) for q in v), ctypes.byref(sb),
                                                   ctypes.byref(vb), ctypes.byref(lb)))
        return EnsembleStats(*(q.value for q in v), sb.value, vb.value, lb.value)
