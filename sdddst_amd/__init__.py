"""sdddst_amd: MI355X-native implicit-step hot path of SDDDST.

The compute path lives in libsdddst_amd.so (hand-written gfx950 HIP kernels
behind the C ABI of include/sdddst_amd.h); this package binds it with ctypes
and provides the seeded workload generator.  Importing the package does not
touch the GPU; creating a Context does.
"""
from ._capi import (  # noqa: F401
    FIELD_EXACT,
    FIELD_FAST,
    FIELD_ORDERED,
    PRECOND_JACOBI,
    PRECOND_LEARNED,
    PRECOND_PAIR,
    PRECOND_STRONG,
    SLOT_BIG_STEP,
    SLOT_DISLOCATIONS,
    SLOT_FIRST_SMALL,
    SLOT_SCRATCH_A,
    SLOT_SCRATCH_B,
    SLOT_SECOND_SMALL,
    SOLVER_DENSE_LU,
    SOLVER_FGMRES,
    SOLVER_HOST,
    V_INIT,
    V_INIT2,
    V_SPEED,
    V_SCRATCH,
    V_SPEED2,
    Context,
    Params,
    SdddstError,
    exported_symbols,
    load_library,
)
from . import configs  # noqa: F401
