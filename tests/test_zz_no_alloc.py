"""sdddst_speeds_noalloc (no counterpart in the reference): the pair speeds
from caller-owned device arrays with the library's static scratch array,
through the C ABI, against the context entry point sdddst_speeds on the same
inputs -- two calls in a row, the second with a smaller input -- and the
size limit refused on the host."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import sdddst_amd as S
from sdddst_amd import configs

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime the library itself links (torch bundles another copy,
    # which cannot share the device with it in one process)
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(x, y, b, sigma, mode):
    # the caller owns inputs and outputs: four device arrays allocated here
    lib = S.load_library()
    hip = _hip()
    n = len(x)
    ptrs = []
    for _ in range(4):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), 8 * n) == 0
        ptrs.append(p)
    try:
        for p, a in zip(ptrs, (x, y, b)):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert hip.hipMemcpy(p, a.ctypes.data, 8 * n, 1) == 0          # host to device
        rc = lib.sdddst_speeds_noalloc(ptrs[0], ptrs[1], ptrs[2], n, sigma, mode, ptrs[3], None)
        assert rc == 0, lib.sdddst_last_error().decode()
        v = np.full(n, np.nan)
        assert hip.hipMemcpy(v.ctypes.data, ptrs[3], 8 * n, 2) == 0        # device to host (synchronous)
        assert hip.hipDeviceSynchronize() == 0
    finally:
        for p in ptrs:
            hip.hipFree(p)
    return v


def _context(x, y, b, sigma, mode):
    with S.Context(len(x), field_mode=mode) as c:
        c.upload_state(0, x, y, b)
        return c.speeds(0, S.V_SCRATCH, sigma, update_tol=False)


@pytest.mark.parametrize("mode", [S.FIELD_FAST, S.FIELD_EXACT])
def test_noalloc_speeds_match_context_entry_point(mode):
    for n, seed, sigma in ((4096, 81, 0.02), (1000, 82, -0.5)):
        print(f"sdddst_speeds_noalloc mode {mode} n {n}", flush=True)
        x, y, b = configs.dislocations(n, seed)
        v = _noalloc(x, y, b, sigma, mode)
        vc = _context(x, y, b, sigma, mode)
        assert np.all(np.isfinite(v))
        strict, ok = O.speed_parity(v, vc, O.speeds_abs(x, y, b), n)
        assert ok.all() and strict.mean() >= 0.995
        assert np.linalg.norm(v - vc) / np.linalg.norm(vc) < 1e-13


def test_noalloc_size_limit_refused():
    lib = S.load_library()
    # refused on the host before any launch (the pointers are never used)
    rc = lib.sdddst_speeds_noalloc(ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 8193, 0.0,
                                   S.FIELD_FAST, ctypes.c_void_p(8), None)
    assert rc == -1 and b"8192" in lib.sdddst_last_error()
    rc = lib.sdddst_speeds_noalloc(None, None, None, 100, 0.0, S.FIELD_FAST, None, None)
    assert rc == -1
