"""Training of the learned preconditioner's network on the GPU.

The training systems are the Newton matrices the simulator assembles in runs
of the constant-stress workload (N dislocations with balanced Burgers
vectors, cutoff multiplier 2, external stress 0.02, precision 1e-6, FGMRES
with the pair-block preconditioner), by default 16 runs at N = 2048 and 4 at
N = 16384 interleaved, 1000 systems each.  After every step attempt the device
context holds the last matrix the simulator assembled, and one loss /
gradient / Adam step of the network runs on it (sdddst_spai_train_step); the
matrices never leave the device.  The learning rate follows a cosine decay
from lr to lr_final over all systems; each run starts a fresh simulation
from its own seed and continues the weights of the previous run.

    python -m sdddst_amd.spai_train --systems 20000 --sizes 2048:16,16384:4 \\
        --out sdddst_amd/data/spai_gnn_h32.f32

writes the weights as raw little-endian float32 values and a JSON record of
the training beside them.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import configs
from ._capi import Context
from .simulation import Simulation

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "spai_gnn_h32.f32")


def save_weights(path: str, w: np.ndarray, record: dict | None = None):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.asarray(w, dtype="<f4").tofile(path)
    if record is not None:
        with open(path + ".json", "w") as f:
            json.dump(record, f, indent=1, sort_keys=True)


def load_weights(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    return np.fromfile(path, dtype="<f4")


def run_sizes(sizes) -> list:
    """[(n, runs), ...] -> the N of each run, the larger sizes spread evenly
    between the smaller ones."""
    total = sum(r for _, r in sizes)
    order = []
    for n, r in sizes:
        order += [((k + 0.5) / r, n) for k in range(r)]
    order.sort()
    assert len(order) == total
    return [n for _, n in order]


def parse_sizes(text: str):
    return [tuple(int(v) for v in part.split(":")) for part in text.split(",")]


def train(systems: int = 20000, sizes=((2048, 16), (16384, 4)), cutoff_multiplier: float = 2.0,
          stress: float = 0.02, probes: int = 8, lr: float = 3e-3, lr_final: float = 2e-4, eps: float = 1e-4,
          init_seed: int = 7, seed0: int = 1000, log=print):
    """Returns (weights float32, record dict)."""
    ns = run_sizes(sizes)
    runs = len(ns)
    per_run = int(math.ceil(systems / runs))
    weights = None
    seen = 0
    t0 = time.time()
    history = []
    for r in range(runs):
        if seen >= systems:
            break
        n = ns[r]
        x, y, b = configs.dislocations(n, seed0 + r)
        sim = Simulation(x, y, b, cutoff_multiplier=cutoff_multiplier, precision=1e-6, initial_step=1e-6,
                         stress="constant", stress_value=stress, solver="fgmres", preconditioner="pair")
        ctx = Context.borrow(sim.device_context, n)
        ctx.spai_train_init(init_seed)
        if weights is not None:
            ctx.spai_set_weights(weights)
        losses = []
        for _ in range(per_run):
            if seen >= systems:
                break
            sim.step()
            frac = seen / max(1, systems - 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
            loss, _ = ctx.spai_train_step(probes, (seed0 + r) * 1_000_003 + seen, lr=rate, eps=eps)
            if not math.isfinite(loss):
                raise FloatingPointError(f"training loss became {loss} at system {seen}")
            losses.append(loss)
            seen += 1
        weights = ctx.spai_get_weights()
        st = sim.state()
        history.append(dict(run=r, n=n, systems=len(losses), loss_mean=float(np.mean(losses)),
                            loss_first=float(losses[0]), loss_last=float(np.mean(losses[-50:])),
                            sim_steps=int(st.succesful_steps), sim_failed=int(st.failed_steps)))
        log(f"run {r} (N={n}): {len(losses)} systems, loss {np.mean(losses):.5f} (last 50 {np.mean(losses[-50:]):.5f}), "
            f"{seen}/{systems} systems, {time.time() - t0:.1f} s")
        ctx.close()
        sim.close()
    record = dict(systems=seen, sizes=[list(s) for s in sizes], runs=len(history), cutoff_multiplier=cutoff_multiplier, stress=stress,
                  probes=probes, lr=lr, lr_final=lr_final, eps=eps, init_seed=init_seed, seed0=seed0,
                  seconds=time.time() - t0, param_count=int(len(weights)), history=history)
    return weights, record


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--systems", type=int, default=20000)
    ap.add_argument("--sizes", default="2048:16,16384:4", help="N:runs,... (runs interleaved)")
    ap.add_argument("--cutoff-multiplier", type=float, default=2.0)
    ap.add_argument("--stress", type=float, default=0.02)
    ap.add_argument("--probes", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    w, rec = train(a.systems, parse_sizes(a.sizes), a.cutoff_multiplier, a.stress, a.probes, a.lr, a.lr_final,
                   init_seed=a.seed, log=lambda s: print(s, flush=True))
    save_weights(a.out, w, rec)
    print(f"wrote {a.out} ({len(w)} weights)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
