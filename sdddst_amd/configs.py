"""Seeded synthetic dislocation configurations (plain-text .dconf compatible).

Positions come from SplitMix64 (Steele, Lea, Flood 2014): each draw
z = mix(state += 0x9E3779B97F4A7C15) gives u = (z >> 11) * 2^-53 in [0, 1),
and a coordinate u - 0.5 in [-0.5, 0.5) (exact in double).  Dislocation k
takes x from draw 2k and y from draw 2k+1 of the seed's stream; the first
half of the indices carry b = +1, the second half b = -1 (balanced +/-1).
Point defects continue the same stream after the dislocations.

The workloads of the benchmark and the tests are named here with their sizes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count successive SplitMix64 outputs (uint64) of the given seed."""
    out = np.empty(count, dtype=np.uint64)
    s = seed & _M64
    for k in range(count):
        s = (s + _GOLDEN) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[k] = z ^ (z >> 31)
    return out


def uniform_coords(seed: int, count: int) -> np.ndarray:
    z = splitmix64(seed, count)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) - 0.5


def dislocations(n: int, seed: int):
    """x, y, b of n dislocations (balanced Burgers vectors)."""
    c = uniform_coords(seed, 2 * n)
    x = c[0::2].copy()
    y = c[1::2].copy()
    b = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    return x, y, b


def point_defects(n_dis: int, pc: int, seed: int):
    """px, py of pc point defects drawn after the n_dis dislocations."""
    c = uniform_coords(seed, 2 * n_dis + 2 * pc)[2 * n_dis:]
    return c[0::2].copy(), c[1::2].copy()


def ensemble_sizes(count: int = 256, seed: int = 3, lo: int = 64, hi: int = 2048):
    """Realization sizes log-uniform in [lo, hi]: round(exp(ln lo + u ln(hi/lo)))."""
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return [int(round(math.exp(math.log(lo) + ui * math.log(hi / lo)))) for ui in u]


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    seed: int
    cutoff_multiplier: float
    precision: float
    stress: str            # "zero", "constant", "fixed-rate"
    stress_value: float    # constant stress or rate
    steps: int
    initial_step: float = 1e-6
    point_defects: int = 0


# The reference's CPU-runnable case and the single-GPU headline case.
SMALL_N256 = Workload("n256-m1-zero-stress", 256, 1, 1.0, 1e-6, "zero", 0.0, 100)
HEADLINE_N16384 = Workload("n16384-m2-const0.02", 16384, 2, 2.0, 1e-6, "constant", 0.02, 200)
LEARNED_N16384 = Workload("n16384-m4-fgmres", 16384, 4, 4.0, 1e-6, "zero", 0.0, 200)
LARGE_N65536 = Workload("n65536-pd2048-m2-ramp", 65536, 5, 2.0, 1e-6, "fixed-rate", 0.0, 100, point_defects=2048)


def write_dconf(path: str, x, y, b):
    """SimulationData::writeDislocationDataToFile (simulation_data.cpp:152-161)."""
    with open(path, "w") as f:
        for xi, yi, bi in zip(x, y, b):
            f.write("%.16e %.16e %.16e\n" % (xi, yi, bi))


def read_dconf(path: str):
    """SimulationData::readDislocationDataFromFile (simulation_data.cpp:119-150)."""
    vals = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return vals[:, 0].copy(), vals[:, 1].copy(), vals[:, 2].copy()


def ensemble_workload(count: int = 256, seed: int = 3, lo: int = 64, hi: int = 2048):
    """The batched-ensemble workload: sizes from ensemble_sizes and, for
    realization r, positions from dislocations(n_r, seed * 1000 + r);
    returns (sizes, x, y, b) with the realizations concatenated."""
    sizes = ensemble_sizes(count, seed, lo, hi)
    xs, ys, bs = [], [], []
    for r, n in enumerate(sizes):
        x, y, b = dislocations(n, seed * 1000 + r)
        xs.append(x)
        ys.append(y)
        bs.append(b)
    return sizes, np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)
