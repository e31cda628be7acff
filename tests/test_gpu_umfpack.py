"""The UMFPACK hand-off of the plain solver path (reference simulation.cpp:83,
386-396): the simulator downloads the CSC into SimulationData's host arrays
and calls umfpack_di_symbolic / numeric / solve / free_* exactly as the
reference does.  SuiteSparse is absent from these machines, so the tests
load a stand-in library with UMFPACK's five entry points (tests/native,
dense LU on the host) through SDDDST_UMFPACK_LIB and check the hand-off: the
matrix the library receives is the oracle's, the call sequence is the
reference's, and a run of the N = 256 configuration matches the oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import configs
from sdddst_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

STUB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libumfpack_stub.so")


@pytest.fixture(scope="module")
def stub():
    if not os.path.exists(STUB):
        O._build(os.path.dirname(STUB), "libumfpack_stub.so")
    os.environ["SDDDST_UMFPACK_LIB"] = STUB
    lib = ctypes.CDLL(STUB)   # the same handle the simulator's dlopen gets
    lib.stub_umfpack_last_matrix.restype = ctypes.c_long
    lib.stub_umfpack_last_matrix.argtypes = [ctypes.POINTER(ctypes.c_int), O.IP, O.IP, O.DP, ctypes.c_long]
    yield lib
    del os.environ["SDDDST_UMFPACK_LIB"]


def _counts(lib):
    c = [ctypes.c_int() for _ in range(4)]
    lib.stub_umfpack_counts(*(ctypes.byref(v) for v in c))
    return [v.value for v in c]


def test_umfpack_handoff_matrix_and_calls(stub):
    n, m = 256, 1.0
    x, y, b = configs.dislocations(n, 1)
    g = Simulation(x, y, b, cutoff_multiplier=m, precision=1e-6, initial_step=1e-6, stress="zero", solver="umfpack",
                   field_mode="exact")
    assert g.solver_name == "umfpack"
    s0 = _counts(stub)
    ok = g.step()
    s1 = _counts(stub)
    # three integrates per step: symbolic + numeric each, ic = 2 solves each,
    # every object freed again (simulation.cpp:83, 386-396)
    assert [s1[0] - s0[0], s1[1] - s0[1], s1[2] - s0[2]] == [3, 3, 6]
    assert s1[3] == 0
    o = O.OracleSim(x, y, b, cutoff_multiplier=m, prec=1e-6, h0=1e-6, stress="zero")
    assert o.step() == ok
    # the last matrix handed over: stage III, assembled at the end of the
    # first small step with h/2
    nn = ctypes.c_int()
    Ap = np.zeros(n + 1, dtype=np.int32)
    nnz = stub.stub_umfpack_last_matrix(ctypes.byref(nn), O.ip(Ap), None, None, 0)
    Ai = np.zeros(nnz, dtype=np.int32)
    Ax = np.zeros(nnz)
    stub.stub_umfpack_last_matrix(ctypes.byref(nn), O.ip(Ap), O.ip(Ai), O.dp(Ax), nnz)
    assert nn.value == n
    rAp, rAi, rAx, _, _, _ = O.jacobian(o.fs_x(), y, b, m, 0.5e-6)
    assert np.array_equal(Ap, rAp) and np.array_equal(Ai, rAi)
    assert np.max(np.abs(Ax - rAx) / np.abs(rAx)) < 1e-12
    assert np.max(np.abs(g.positions() - o.x())) < 1e-12
    g.close()
    o.close()


def test_umfpack_run_config1_matches_oracle(stub):
    # N = 256, m = 1, zero stress, precision 1e-6: 100 successful steps with
    # the plain solver path, against the oracle's run
    x, y, b = configs.dislocations(256, 1)
    o = O.OracleSim(x, y, b, cutoff_multiplier=1.0, prec=1e-6, h0=1e-6, stress="zero")
    while o.succesful_steps < 100:
        o.step()
    g = Simulation(x, y, b, cutoff_multiplier=1.0, precision=1e-6, initial_step=1e-6, stress="zero", solver="auto",
                   field_mode="exact")
    assert g.solver_name == "umfpack"      # auto resolves to UMFPACK when it loads
    g.run(100)
    st = g.state()
    assert st.succesful_steps == 100 and st.failed_steps == o.failed_steps
    assert abs(st.sim_time - o.time) <= 1e-9 * o.time
    assert np.max(np.abs(g.positions() - o.x())) < 1e-8
    assert _counts(stub)[3] == 0
    g.close()
    o.close()


def test_umfpack_missing_library_is_an_error(stub, tmp_path):
    os.environ["SDDDST_UMFPACK_LIB"] = str(tmp_path / "no-such-libumfpack.so")
    try:
        x, y, b = configs.dislocations(64, 3)
        from sdddst_amd._capi import SdddstError
        with pytest.raises(SdddstError, match="UMFPACK"):
            Simulation(x, y, b, solver="umfpack")
    finally:
        os.environ["SDDDST_UMFPACK_LIB"] = STUB
