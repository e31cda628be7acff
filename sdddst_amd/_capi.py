"""ctypes binding of the C ABI in include/sdddst_amd.h (libsdddst_amd.so).

The library is loaded from this package directory; there is no fallback: if
it is missing or the device is absent the calls raise.  Array arguments are
numpy float64/int32 arrays; every call checks the returned status and raises
SdddstError with the library's message.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdddst_amd.so")

OK = 0
SLOT_DISLOCATIONS, SLOT_BIG_STEP, SLOT_FIRST_SMALL, SLOT_SECOND_SMALL, SLOT_SCRATCH_A, SLOT_SCRATCH_B = range(6)
V_INIT, V_INIT2, V_SPEED, V_SPEED2, V_SCRATCH = range(5)
FIELD_FAST, FIELD_EXACT, FIELD_ORDERED = 0, 1, 2
SOLVER_DENSE_LU, SOLVER_FGMRES, SOLVER_HOST = 0, 1, 2
PRECOND_JACOBI, PRECOND_PAIR, PRECOND_LEARNED, PRECOND_STRONG = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class SdddstError(RuntimeError):
    """A non-zero status from the C ABI."""


class Params(ctypes.Structure):
    _fields_ = [
        ("cutoff_multiplier", ctypes.c_double),
        ("precision", ctypes.c_double),
        ("A", ctypes.c_double),
        ("KASQR", ctypes.c_double),
        ("iteration_count", ctypes.c_int),
        ("field_mode", ctypes.c_int),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "sdddst_version": (ctypes.c_char_p, []),
    "sdddst_last_error": (ctypes.c_char_p, []),
    "sdddst_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "sdddst_ctx_destroy": (ctypes.c_int, [_vp]),
    "sdddst_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(Params)]),
    "sdddst_get_cutoff": (ctypes.c_int, [_vp, _dp, _dp]),
    "sdddst_upload_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    "sdddst_download_x": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int]),
    "sdddst_copy_slot": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "sdddst_swap_slots": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "sdddst_upload_points": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int]),
    "sdddst_speeds": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
    "sdddst_get_tolerance": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "sdddst_reset_tolerance": (ctypes.c_int, [_vp]),
    "sdddst_download_speed": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int]),
    "sdddst_upload_speed": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int]),
    "sdddst_copy_speed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "sdddst_jacobian_csc": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _ip, _dp, _dp, _ip, _i64p],
    ),
    "sdddst_residual": (
        ctypes.c_int,
        [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    ),
    "sdddst_solver_select": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "sdddst_set_preconditioner": (ctypes.c_int, [_vp, ctypes.c_int]),
    "sdddst_solver_factor": (ctypes.c_int, [_vp]),
    "sdddst_solver_solve": (ctypes.c_int, [_vp, _dp]),
    "sdddst_solver_stats": (ctypes.c_int, [_vp, _ip, _dp, _ip, _i64p, _i64p]),
    "sdddst_solver_precond_fallbacks": (ctypes.c_int, [_vp, _ip]),
    "sdddst_solver_history": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _ip, _ip]),
    "sdddst_speeds_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp]),
    "sdddst_fp64_probe": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
    "sdddst_solver_set_direct_limit": (ctypes.c_int, [_vp, ctypes.c_int]),
    "sdddst_set_learned_gate": (ctypes.c_int, [_vp, ctypes.c_int]),
    "sdddst_spai_param_count": (ctypes.c_int, []),
    "sdddst_spai_set_weights": (ctypes.c_int, [_vp, _fp, ctypes.c_int]),
    "sdddst_spai_get_weights": (ctypes.c_int, [_vp, _fp, ctypes.c_int]),
    "sdddst_spai_train_init": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "sdddst_spai_train_step": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _dp, _dp],
    ),
    "sdddst_spai_graph": (ctypes.c_int, [_vp, _i64p, _ip, _ip, _dp, _dp, _fp, _fp, _dp, _dp, _ip]),
    "sdddst_newton_update": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "sdddst_error_control": (
        ctypes.c_int,
        [_vp, ctypes.c_double, ctypes.c_double, _ip, _dp, _dp, _ip],
    ),
    "sdddst_speed_sums": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    "sdddst_strain_increment": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "sdddst_stage_end": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                        ctypes.POINTER(ctypes.c_int), _dp, _dp, ctypes.POINTER(ctypes.c_int), _dp, _dp]),
    "sdddst_synchronize": (ctypes.c_int, [_vp]),
    "sdddst_field_eval": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp]),
    "sdddst_libm_eval": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, _dp, _dp]),
    "sdddst_ev_decompose": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    "sdddst_ev_analysis": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sdddst_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "sdddst_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_char_p, _ip, _dp]),
}

_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libsdddst_amd.so and declare every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SdddstError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__; __graft_entry__.build()'` "
            "or `make -C sdddst_amd`"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Every entry point bound from Python (this module and ensemble.py)."""
    from .ensemble import _SIGS as ensemble_sigs
    return list(_SIGNATURES) + list(ensemble_sigs)


def _dptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_dp)


def _iptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_ip)


def _fptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_fp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context (sdddst_ctx) holding the state of a system of n
    dislocations and pc point defects."""

    def __init__(self, n: int, pc: int = 0, cutoff_multiplier: float = 1.0, precision: float = 1e-6,
                 A: float = 1e-4 * 16.0, KASQR: float = 1.65 * 1.65 * 1e6 / 256.0, iteration_count: int = 2,
                 field_mode: int = FIELD_FAST):
        self.lib = load_library()
        self.n, self.pc = int(n), int(pc)
        self.params = Params(cutoff_multiplier, precision, A, KASQR, iteration_count, field_mode)
        h = _vp()
        self._check(self.lib.sdddst_ctx_create(self.n, self.pc, ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    # -- plumbing --------------------------------------------------------
    def _check(self, rc: int):
        if rc != OK:
            msg = self.lib.sdddst_last_error().decode(errors="replace")
            raise SdddstError(f"status {rc}: {msg}")

    @classmethod
    def borrow(cls, handle, n: int, pc: int = 0) -> "Context":
        """Non-owning view of a context created elsewhere (the simulator's,
        Simulation.device_context()); closing it leaves the context alive."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.n, self.pc = int(n), int(pc)
        self.params = None
        self.h = handle
        self._borrowed = True
        return self

    def close(self):
        if getattr(self, "h", None):
            if not getattr(self, "_borrowed", False):
                self.lib.sdddst_ctx_destroy(self.h)
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

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self.lib.sdddst_set_params(self.h, ctypes.byref(self.params)))

    def cutoff(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.sdddst_get_cutoff(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # -- state -----------------------------------------------------------
    def upload_state(self, slot: int, x, y=None, b=None):
        x = _f64(x)
        y = None if y is None else _f64(y)
        b = None if b is None else _f64(b)
        self._check(self.lib.sdddst_upload_state(self.h, slot, _dptr(x), _dptr(y), _dptr(b), self.n))

    def download_x(self, slot: int) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self.lib.sdddst_download_x(self.h, slot, _dptr(out), self.n))
        return out

    def copy_slot(self, dst: int, src: int):
        self._check(self.lib.sdddst_copy_slot(self.h, dst, src))

    def swap_slots(self, a: int, b: int):
        self._check(self.lib.sdddst_swap_slots(self.h, a, b))

    def upload_points(self, px, py):
        px, py = _f64(px), _f64(py)
        self._check(self.lib.sdddst_upload_points(self.h, _dptr(px), _dptr(py), len(px)))

    # -- K1 --------------------------------------------------------------
    def speeds(self, cfg_slot: int = 0, out_buf: int = V_INIT, sigma: float = 0.0, update_tol: bool = True,
               download: bool = True) -> Optional[np.ndarray]:
        out = np.empty(self.n) if download else None
        self._check(self.lib.sdddst_speeds(self.h, cfg_slot, out_buf, sigma, int(update_tol), _dptr(out)))
        return out

    def tolerance(self) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self.lib.sdddst_get_tolerance(self.h, _dptr(out), self.n))
        return out

    def reset_tolerance(self):
        self._check(self.lib.sdddst_reset_tolerance(self.h))

    def download_speed(self, buf: int) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self.lib.sdddst_download_speed(self.h, buf, _dptr(out), self.n))
        return out

    def upload_speed(self, buf: int, v):
        v = _f64(v)
        self._check(self.lib.sdddst_upload_speed(self.h, buf, _dptr(v), self.n))

    def copy_speed(self, dst: int, src: int):
        self._check(self.lib.sdddst_copy_speed(self.h, dst, src))

    # -- K2 --------------------------------------------------------------
    def jacobian_nnz(self, cfg_slot: int, h: float) -> int:
        nnz = ctypes.c_int64()
        self._check(self.lib.sdddst_jacobian_csc(self.h, cfg_slot, h, 1, None, None, None, None, None,
                                                 ctypes.byref(nnz)))
        return nnz.value

    def jacobian_csc(self, cfg_slot: int, h: float, rebuild: bool = True):
        """Build and download (Ap, Ai, Ax, dVec, diag)."""
        nnz = self.jacobian_nnz(cfg_slot, h) if rebuild else None
        if nnz is None:
            tmp = ctypes.c_int64()
            self._check(self.lib.sdddst_jacobian_csc(self.h, cfg_slot, h, 0, None, None, None, None, None,
                                                     ctypes.byref(tmp)))
            nnz = tmp.value
        Ap = np.empty(self.n + 1, dtype=np.int32)
        Ai = np.empty(nnz, dtype=np.int32)
        Ax = np.empty(nnz)
        dv = np.empty(self.n)
        dg = np.empty(self.n, dtype=np.int32)
        out_nnz = ctypes.c_int64()
        self._check(self.lib.sdddst_jacobian_csc(self.h, cfg_slot, h, 0, _iptr(Ap), _iptr(Ai), _dptr(Ax), _dptr(dv),
                                                 _iptr(dg), ctypes.byref(out_nnz)))
        return Ap, Ai, Ax, dv, dg

    # -- K3/K4 and solver ------------------------------------------------
    def residual(self, h: float, new_slot: int, old_slot: int, csp: int, isp: int, download: bool = False):
        out = np.empty(self.n) if download else None
        self._check(self.lib.sdddst_residual(self.h, h, new_slot, old_slot, csp, isp, _dptr(out)))
        return out

    def solver_select(self, solver: int, rel_tol: float = 0.0, max_iter: int = 0, restart: int = 0):
        self._check(self.lib.sdddst_solver_select(self.h, solver, rel_tol, max_iter, restart))

    def set_preconditioner(self, precond: int):
        self._check(self.lib.sdddst_set_preconditioner(self.h, precond))

    def set_learned_gate(self, max_strong_rows: int):
        """Most strongly shifted diagonals (|A_jj - 1| > 0.25) a matrix may
        have for the learned preconditioner (-1: always; -2: max(16, N/256))."""
        self._check(self.lib.sdddst_set_learned_gate(self.h, max_strong_rows))

    def solver_factor(self):
        self._check(self.lib.sdddst_solver_factor(self.h))

    def solver_solve(self, download: bool = False):
        out = np.empty(self.n) if download else None
        self._check(self.lib.sdddst_solver_solve(self.h, _dptr(out)))
        return out

    def solver_stats(self):
        it, rr, fb = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.sdddst_solver_stats(self.h, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(fb),
                                                 None, None))
        return it.value, rr.value, fb.value

    def precond_fallbacks(self) -> int:
        v = ctypes.c_int()
        self._check(self.lib.sdddst_solver_precond_fallbacks(self.h, ctypes.byref(v)))
        return v.value

    def solver_history(self, max_records: int = 1024):
        """Recent iterative-solver attempts as an array of rows {precond,
        iterations, relative residual, ||g||, outcome, h, strong-block
        rows} (clears the log),
        and the count of solves retried with the strong-row block."""
        return solver_history(self.h, max_records)

    def set_direct_limit(self, max_n: int):
        self._check(self.lib.sdddst_solver_set_direct_limit(self.h, max_n))

    # -- learned preconditioner ------------------------------------------
    def spai_param_count(self) -> int:
        return int(self.lib.sdddst_spai_param_count())

    def spai_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.sdddst_spai_set_weights(self.h, _fptr(w), len(w)))

    def spai_get_weights(self) -> np.ndarray:
        w = np.empty(self.spai_param_count(), dtype=np.float32)
        self._check(self.lib.sdddst_spai_get_weights(self.h, _fptr(w), len(w)))
        return w

    def spai_train_init(self, seed: int):
        self._check(self.lib.sdddst_spai_train_init(self.h, seed))

    def spai_train_step(self, probes: int, seed: int, lr: float = 0.0, eps: float = 1e-4, want_grad: bool = False):
        """Loss and gradient on the current Jacobian, then one Adam step
        when lr > 0: returns (loss, grad or None)."""
        loss = ctypes.c_double()
        g = np.empty(self.spai_param_count()) if want_grad else None
        self._check(self.lib.sdddst_spai_train_step(self.h, probes, seed, eps, lr, ctypes.byref(loss), _dptr(g)))
        return loss.value, g

    def spai_graph(self):
        """Last evaluated core graph and preconditioner values as a dict."""
        ne = ctypes.c_int64()
        self._check(self.lib.sdddst_spai_graph(self.h, ctypes.byref(ne), None, None, None, None, None, None, None,
                                               None, None))
        n, m = self.n, ne.value
        out = dict(Cp=np.empty(n + 1, np.int32), Ci=np.empty(m, np.int32), Ce=np.empty(m), Cd=np.empty(n),
                   gamma_e=np.empty(m, np.float32), gamma_d=np.empty(n, np.float32), bd=np.empty(n), bo=np.empty(n),
                   bcol=np.empty(n, np.int32))
        self._check(self.lib.sdddst_spai_graph(
            self.h, ctypes.byref(ne), _iptr(out["Cp"]), _iptr(out["Ci"]), _dptr(out["Ce"]), _dptr(out["Cd"]),
            _fptr(out["gamma_e"]), _fptr(out["gamma_d"]), _dptr(out["bd"]), _dptr(out["bo"]), _iptr(out["bcol"])))
        return out

    def newton_update(self, slot: int, delta=None):
        d = None if delta is None else _f64(delta)
        self._check(self.lib.sdddst_newton_update(self.h, slot, _dptr(d)))

    def error_control(self, h: float, hmax: float = -1.0):
        acc, hn, mr, sel = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.sdddst_error_control(self.h, h, hmax, ctypes.byref(acc), ctypes.byref(hn),
                                                  ctypes.byref(mr), ctypes.byref(sel)))
        return bool(acc.value), hn.value, mr.value, sel.value

    def speed_sums(self, buf: int):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.sdddst_speed_sums(self.h, buf, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def strain_increment(self, old_slot: int, new_slot: int) -> float:
        r = ctypes.c_double()
        self._check(self.lib.sdddst_strain_increment(self.h, old_slot, new_slot, ctypes.byref(r)))
        return r.value

    def synchronize(self):
        self._check(self.lib.sdddst_synchronize(self.h))

    # -- per-pair evaluation and timing ---------------------------------
    def field_eval(self, which: int, dx, dy) -> np.ndarray:
        dx, dy = _f64(dx), _f64(dy)
        out = np.empty_like(dx)
        self._check(self.lib.sdddst_field_eval(self.h, which, len(dx), _dptr(dx), _dptr(dy), _dptr(out)))
        return out

    def libm_eval(self, fn: int, x) -> np.ndarray:
        x = _f64(x)
        out = np.empty_like(x)
        self._check(self.lib.sdddst_libm_eval(self.h, fn, len(x), _dptr(x), _dptr(out)))
        return out

    def ev_decompose(self, slot: int, vectors: bool = True, speeds=None, stress_rate: float = 1.0) -> dict:
        """EV analysis of the positions in slot (SpectralDecompositor::decompose
        and the per-mode columns of TimeSeriesProcessor::run): ascending
        eigenvalues, eigenvectors as rows (None unless vectors), participation
        numbers, s_i = sum_j v_ij b_j and s_i^2 / lambda_i; with speeds, also
        the eigenvalue derivatives over stress_rate ("deriv")."""
        n = self.n
        lam, pn, gan, sb = np.empty(n), np.empty(n), np.empty(n), np.empty(n)
        vec = np.empty((n, n)) if vectors else None
        if speeds is None:
            self._check(self.lib.sdddst_ev_decompose(self.h, slot, _dptr(lam), _dptr(vec) if vectors else None,
                                                     _dptr(pn), _dptr(gan), _dptr(sb)))
            return {"eigenvalues": lam, "eigenvectors": vec, "pn": pn, "gan": gan, "sum_b": sb}
        sp = np.ascontiguousarray(speeds, dtype=np.float64)
        if sp.shape != (n,):
            raise ValueError(f"speeds must have {n} entries")
        deriv = np.empty(n)
        self._check(self.lib.sdddst_ev_analysis(self.h, slot, _dptr(sp), float(stress_rate), _dptr(lam),
                                                _dptr(vec) if vectors else None, _dptr(pn), _dptr(gan), _dptr(sb),
                                                _dptr(deriv)))
        return {"eigenvalues": lam, "eigenvectors": vec, "pn": pn, "gan": gan, "sum_b": sb, "deriv": deriv}

    def timing(self, on: bool, pair_only: bool = False):
        timing(self.h, on, pair_only)

    def kernel_stats(self, kernel: str):
        return kernel_stats(self.h, kernel)


def timing(ctx_handle, on, pair_only: bool = False):
    """Start (clearing the counts) or stop the event timers; pair_only
    records only the pair kernel's events."""
    lib = load_library()
    rc = lib.sdddst_timing(ctx_handle, (2 if pair_only else 1) if on else 0)
    if rc != OK:
        raise SdddstError(lib.sdddst_last_error().decode(errors="replace"))


def kernel_stats(ctx_handle, kernel: str):
    """(launch count, total ms) of a kernel kind since timing was enabled."""
    lib = load_library()
    cnt, ms = ctypes.c_int(), ctypes.c_double()
    rc = lib.sdddst_kernel_stats(ctx_handle, kernel.encode(), ctypes.byref(cnt), ctypes.byref(ms))
    if rc != OK:
        raise SdddstError(lib.sdddst_last_error().decode(errors="replace"))
    return cnt.value, ms.value



def fp64_probe(iters: int = 20000):
    """sdddst_fp64_probe: (FMA-stream TFLOP/s, FMA and multiply wave-level
    VALU instructions per second) sustained by this device."""
    lib = load_library()
    t, f, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib.sdddst_fp64_probe(int(iters), ctypes.byref(t), ctypes.byref(f), ctypes.byref(m))
    if rc != OK:
        raise SdddstError(lib.sdddst_last_error().decode(errors="replace"))
    return t.value, f.value, m.value


def solver_history(ctx_handle, max_records: int = 1024):
    """sdddst_solver_history on a context handle: (records [k, 7], strong-row retries)."""
    lib = load_library()
    out = np.empty((max_records, 7))
    k, ext = ctypes.c_int(), ctypes.c_int()
    rc = lib.sdddst_solver_history(ctx_handle, _dptr(out), max_records, ctypes.byref(k), ctypes.byref(ext))
    if rc != OK:
        raise SdddstError(lib.sdddst_last_error().decode(errors="replace"))
    return out[:k.value].copy(), ext.value
