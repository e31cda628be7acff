"""Write the golden vectors of tests/golden/ from the reference's own
translation units (oracle/_ref/libsdddst_ref.so, built by `make -C oracle`
from the reference sources): AnalyticField::xy / xy_diff_x / xy_diff_x2 at
fixed points,
the reference's pair loop (simulation.cpp:86-131 order over
AnalyticField::xy) at N=256, PrecisionHandler sequences and the stress
protocols.  Inputs are seeded; outputs are stored bit-exactly in .npz.

Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from sdddst_amd import configs  # noqa: E402


def field_points():
    rng = np.random.default_rng(20240601)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 3000), rng.uniform(-1.5e-3, 1.5e-3, 800),
                         rng.uniform(-0.06, 0.06, 600), rng.uniform(-0.5, 0.5, 200),
                         np.array([0.0, -0.5, 0.49999999999999994, 0.25, -0.25, 1e-9, -1e-9, 7e-4, -7e-4, 0.1])])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 3000), rng.uniform(-1.5e-3, 1.5e-3, 800),
                         rng.uniform(-0.06, 0.06, 600), np.zeros(200),
                         np.array([0.3, 0.1, -0.2, 0.5, -0.5, 2e-4, -3e-4, 0.0, 7e-4, 0.0])])
    return dx, dy


def main():
    ref = O.reference()
    if ref is None:
        sys.exit("oracle/_ref/libsdddst_ref.so is missing: run `make -C oracle ref` where the reference is present")
    dx, dy = field_points()
    xy = np.empty_like(dx)
    xyd = np.empty_like(dx)
    ref.ref_xy_batch(len(dx), O.dp(dx), O.dp(dy), O.dp(xy), O.dp(xyd))
    np.savez(os.path.join(HERE, "field_points.npz"), dx=dx, dy=dy, xy=xy, xy_diff_x=xyd)
    # AnalyticField::xy_diff_x2 (the EV analysis' eigenvalue derivatives)
    xyd2 = np.array([ref.ref_xy_diff_x2(float(u), float(v)) for u, v in zip(dx, dy)])
    np.savez(os.path.join(HERE, "field_points_dx2.npz"), dx=dx, dy=dy, xy_diff_x2=xyd2)

    n = 256
    x, y, b = configs.dislocations(n, 1)
    v = np.empty(n)
    ref.ref_speeds(n, O.dp(x), O.dp(y), O.dp(b), 0.0, O.dp(v))
    vs = np.empty(n)
    ref.ref_speeds(n, O.dp(x), O.dp(y), O.dp(b), 0.02, O.dp(vs))
    np.savez(os.path.join(HERE, "speeds_n256_seed1.npz"), x=x, y=y, b=b, v_zero=v, v_sigma002=vs)

    rng = np.random.default_rng(77)
    cases = []
    for case in range(40):
        size = int(rng.integers(2, 50))
        prec = float(10.0 ** rng.uniform(-8, -4))
        nupd = int(rng.integers(0, 200))
        r2 = rng.uniform(0, 1e-6, nupd) * (10.0 ** rng.uniform(-6, 0, nupd))
        if case % 5 == 0 and nupd:
            r2[rng.integers(0, nupd)] = 0.0     # coincident pair: tolerance EPS
        ids = rng.integers(0, size, nupd).astype(np.uint32)
        nerr = size
        err = np.abs(rng.normal(0, 10.0 ** rng.uniform(-9, -5), nerr))
        err[rng.integers(0, nerr)] = 0.0
        eids = np.arange(size, dtype=np.uint32)
        old = float(10.0 ** rng.uniform(-7, -3))
        newstep = ctypes.c_double()
        mr = ref.ref_precision_sequence(size, prec, nupd, O.dp(r2), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                        nerr, O.dp(err), eids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), old,
                                        ctypes.byref(newstep))
        cases.append((size, prec, r2, ids, err, old, mr, newstep.value))
    np.savez(os.path.join(HERE, "precision_handler.npz"),
             size=np.array([c[0] for c in cases]), prec=np.array([c[1] for c in cases]),
             r2=np.concatenate([c[2] for c in cases]), r2_len=np.array([len(c[2]) for c in cases]),
             ids=np.concatenate([c[3] for c in cases]),
             err=np.concatenate([c[4] for c in cases]), err_len=np.array([len(c[4]) for c in cases]),
             old=np.array([c[5] for c in cases]), max_ratio=np.array([c[6] for c in cases]),
             new_step=np.array([c[7] for c in cases]))

    rows = []
    for kind in (0, 1, 2):
        for t in (0.0, 1e-6, 0.37, 12.5):
            for strain in (0.0, 1e-7, 5.0):
                for typ in range(4):
                    rows.append((kind, 0.75, 2.5, t, strain, typ,
                                 ref.ref_protocol_stress(kind, 0.75, 2.5, t, strain, typ)))
    np.savez(os.path.join(HERE, "stress_protocols.npz"), table=np.array(rows))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
