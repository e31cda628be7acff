#!/usr/bin/env python3
"""Benchmark: successful implicit steps/s of the SDDDST implicit-step hot path
on one MI355X, with the pair-speed kernel's FP64 roofline fraction and a
bounded CPU baseline.

Workload (default): N = 16384 edge dislocations, positions uniform in the
unit periodic cell from seed 2, balanced +/-1 Burgers vectors, constant
external stress 0.02, cutoff multiplier 2, precision 1e-6, initial step
1e-6.  A "step" is one successful implicit step of sdddstCore::Simulation
(three stages, Jacobian assemblies, Newton solves, the error controller and
the log); failed attempts inside the timed region count against the rate.
The headline solves with FGMRES and the learned preconditioner; the same
run with the pair-block preconditioner alone (the method without the
network), a few steps of the direct-solver path and the batched ensemble of
256 realizations (N 64..2048, 500 steps each) are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precond learned]
Prints one JSON line.  `python bench.py --copy-calibration` runs only a plain
device copy of known size (the reference kernel for scaling the HBM byte
counters of a profiler run); `python bench.py --ev-analysis N` times only the
EV analysis of one N-dislocation snapshot.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "successful implicit steps/s at N=16384, 1e-6 precision; GPU pairs/s vs FP64 peak"
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector, AMD public specification
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR = 745          # cost-model FP64 operations of one tau_xy pair term


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--solver", default="fgmres", choices=["fgmres", "dense-lu", "auto", "umfpack"])
    p.add_argument("--field-mode", default="fast", choices=["fast", "exact"])
    p.add_argument("--precond", default="learned", choices=["pair", "jacobi", "learned"])
    p.add_argument("--compare-steps", type=int, default=None,
                   help="steps of the same workload with the pair-block preconditioner, from t = 0 like the "
                        "headline (default: --steps, so both rates cover the same window; 0: skip)")
    p.add_argument("--ensemble-steps", type=int, default=500,
                   help="steps per realization of the 256-realization ensemble (0: skip)")
    p.add_argument("--workload", default="n16384-m2-const0.02", choices=["n16384-m2-const0.02",
                                                                          "n16384-m4-zero-learned",
                                                                          "n65536-pd2048-m2-ramp"],
                   help="the headline workload, or the large combined run (N=65536 + 2048 point defects, "
                        "stress ramped at 4e6 per unit time so that it reaches about 0.1 at step 100, "
                        "learned preconditioner, 100 steps from t = 0, no warm-up)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a 512 MiB device-to-device copy kernel and exit")
    p.add_argument("--plain-steps", type=int, default=5,
                   help="also time this many steps of the direct-solver path (0: skip)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--assembly-reps", type=int, default=10,
                   help="full Jacobian assemblies of the headline positions timed on their own (0: skip)")
    p.add_argument("--assembly-only", action="store_true",
                   help="run only the Jacobian assembly leg and print its record")
    p.add_argument("--cpu-seconds", type=float, default=8.0,
                   help="seconds of the pair-speed row sample of the CPU baseline")
    p.add_argument("--ev-analysis", type=int, metavar="N", default=0,
                   help="only time the EV analysis of one N-dislocation snapshot (Hessian, eigen-decomposition, "
                        "observables, eigenvalue derivatives) and print its own JSON line")
    p.add_argument("--trace-gaps", metavar="DIR",
                   help="no GPU work: summarise the device idle time of a rocprofv3 --kernel-trace CSV "
                        "output directory (gaps between kernels over the last 60 %% of the trace)")
    a = p.parse_args()
    if a.compare_steps is None:
        a.compare_steps = a.steps
    return a


def trace_gaps(d):
    """Device idle time between kernels in a rocprofv3 kernel trace: the
    union of kernel intervals over the last 60 % of the dispatches (the
    timed steps of a bench run) against their span, and the gaps grouped by
    the kernel that ends before them."""
    import collections
    import csv
    import glob
    import re
    f = glob.glob(os.path.join(d, "*kernel_trace.csv"))[0]
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in csv.DictReader(open(f)))
    t0 = iv[int(len(iv) * 0.4)][0]

    def short(name):
        m = re.search(r"(k_\w+|__amd_rocclr_\w+|rocprim)", name)
        return m.group(1) if m else name[:40]

    busy, gaps = 0, collections.defaultdict(lambda: [0, 0])
    cur_s = cur_e = None
    last = ""
    for st, en, nm in iv:
        if st < t0:
            continue
        if cur_e is None:
            cur_s, cur_e = st, en
        elif st > cur_e:
            busy += cur_e - cur_s
            g = gaps[short(last)]
            g[0] += 1
            g[1] += st - cur_e
            cur_s, cur_e = st, en
        else:
            cur_e = max(cur_e, en)
        last = nm if en >= cur_e else last
    busy += cur_e - cur_s
    span = iv[-1][1] - t0
    print(f"span {span / 1e6:.2f} ms, kernels busy {busy / 1e6:.2f} ms, idle {(span - busy) / 1e6:.2f} ms "
          f"({100 * (span - busy) / span:.1f} %)")
    print("idle after kernel: count, total ms, mean us")
    for k, (c, t) in sorted(gaps.items(), key=lambda kv: -kv[1][1])[:15]:
        print(f"  {k:28s} {c:6d} {t / 1e6:8.2f} {t / c / 1e3:8.2f}")


def make_sim(args, solver, x, y, b, precond=None):
    from sdddst_amd.simulation import Simulation
    return Simulation(x, y, b, cutoff_multiplier=2.0, precision=1e-6, initial_step=1e-6, stress="constant",
                      stress_value=0.02, calculate_strain=True, solver=solver, field_mode=args.field_mode,
                      preconditioner=precond or args.precond,
                      solver_tol=1e-10, solver_max_iter=400, solver_restart=40)


def copy_calibration():
    """A plain copy of 512 MiB (64 Mi float64) through torch: 512 MiB read,
    512 MiB written; the profiler's FETCH_SIZE / WRITE_SIZE of this dispatch
    give the scale of those counters."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": c.numel() * 8, "dispatches": 3}))


def traffic_record():
    """HBM bytes per launch of the pair kernel from the committed profiler
    summary (two counter runs, scaled by the copy calibration)."""
    path = os.path.join(HERE, "profiles", "pair_kernel_n16384_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec.get("hbm_bytes_per_launch"), rec
    except (OSError, ValueError):
        return None, None


def valu_record():
    """Executed instructions per pair of the pair kernel from the committed
    counter profile (SQ_INSTS_VALU and the FP64 ADD/MUL/FMA/TRANS counts of
    one rocprofv3 run over the benchmark workload)."""
    path = os.path.join(HERE, "profiles", "pair_kernel_n16384_valu.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def pair_roofline(pairs, avg_ms, traffic, traffic_rec):
    """The pair kernel's roofline record.  The kernel is bound by the FP64
    VALU pipe (arithmetic intensity about N * 160 / 24 flop/byte), so
    `achieved` counts the FP64 flops the kernel executes per launch (the
    counter profile's ADD + MUL + 2 FMA + TRANS FP64 instructions, per pair
    of the workload: the kernel evaluates the pairs of the tiles that are
    not x-separated, the series kernels sum the others) over its average
    launch time, against the MI355X FP64 vector peak.  The reference expression tree's cost model (745 flops per
    pair: 165 add/sub/mul + 10 div x 15 + 10 cosh x 40 + 1 cos x 30) is kept
    beside it as `cost_model_tflops`: the production field needs far fewer
    operations per pair than that tree, so the cost model's rate passes the
    peak and is a pair rate, not a hardware figure.  `valu_issue_frac` is
    VALU instructions per second against the issue rate of one 64-lane
    instruction per CU per clock (= FP64 peak / 2 lane-ops/s)."""
    rec = valu_record()
    t = avg_ms * 1e-3
    out = {"bound": "fp64-valu", "kernel": "k_pair_tiles", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
           "achieved": None, "frac": None, "traffic": traffic, "pairs_per_launch": pairs, "avg_launch_ms": avg_ms,
           "cost_model_flops_per_pair": FLOPS_PER_PAIR,
           "cost_model_tflops": FLOPS_PER_PAIR * pairs / t / 1e12,
           "cost_model": "745 FP64 flops per pair: 165 add/sub/mul + 10 div x 15 + 10 cosh x 40 + 1 cos x 30"}
    if traffic_rec:
        out["traffic_source"] = traffic_rec.get("source")
    if rec:
        flops = rec["fp64_flops_per_pair"]
        insts = rec["valu_insts_per_pair"]
        achieved = flops * pairs / t / 1e12
        out.update({"achieved": achieved, "frac": achieved / FP64_PEAK_TFLOPS,
                    "executed_fp64_flops_per_pair": flops, "valu_insts_per_pair": insts,
                    "valu_issue_frac": insts * pairs / t / (FP64_PEAK_TFLOPS * 1e12 / 2),
                    "execution_source": rec.get("source")})
    return out


def issue_probe(roof):
    """The FP64 VALU rate this device sustains on independent FMA and
    multiply streams (sdddst_fp64_probe), measured after the timed run, and
    the pair kernel's own VALU instruction rate against it: the published
    78.6 TFLOP/s assumes the 2.4 GHz peak clock, which FP64-heavy kernels do
    not hold."""
    from sdddst_amd._capi import fp64_probe
    tf, fma_rate, mul_rate = fp64_probe()
    rec = {"fma_stream_tflops": tf, "fma_stream_valu_per_s": fma_rate, "mul_stream_valu_per_s": mul_rate,
           "source": "sdddst_fp64_probe: 8 independent chains per lane, 16 waves per CU, best of 3 launches"}
    insts = roof.get("valu_insts_per_pair")
    if insts and roof.get("avg_launch_ms"):
        k = insts * roof["pairs_per_launch"] / 64.0 / (roof["avg_launch_ms"] * 1e-3)
        rec.update({"pair_kernel_valu_per_s": k, "pair_kernel_frac_of_fma_stream": k / fma_rate,
                    "executed_fp64_frac_of_fma_stream": roof["achieved"] / tf if roof.get("achieved") else None})
    return rec


def assembly_run(n=16384, m=2.0, reps=10):
    """One full Jacobian assembly (calculateJacobian, simulation.cpp:236-384:
    count, scan, fill, values, column sums, d, scaling) of the headline
    workload's initial positions, timed with HIP events on the context's
    stream around each assembly (host read-back of nnz included), and
    compared with the compulsory bytes 12 nnz + 40 N + 4 (write Ap, Ai, Ax,
    dVec and the diagonal positions; read x, y, b).  The HBM bytes the
    kernels actually move come from the committed counter profile."""
    import sdddst_amd as S
    from sdddst_amd import configs
    x, y, b = configs.dislocations(n, 2)
    with S.Context(n, cutoff_multiplier=m, field_mode=S.FIELD_FAST) as c:
        c.upload_state(0, x, y, b)
        nnz = c.jacobian_nnz(0, 1e-6)
        c.synchronize()
        c.timing(True)
        for _ in range(reps):
            c.jacobian_nnz(0, 1e-6)
        c.synchronize()
        cnt, ms = c.kernel_stats("jacobian")
        c.timing(False)
    avg = ms / max(cnt, 1)
    alg = 12 * nnz + 40 * n + 4
    rec = {"workload": f"N={n}, m={m:g}, seed 2 positions, fast field, full rebuild", "nnz": nnz,
           "algorithmic_bytes": alg, "avg_assembly_ms": avg, "assemblies": cnt,
           "achieved_gbs": alg / (avg * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
           "frac": alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS}
    path = os.path.join(HERE, "profiles", "jacobian_n16384_m2_traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)
        rec.update({"hbm_bytes_per_assembly": prof["hbm_bytes_per_assembly"],
                    "kernel_busy_ms": prof.get("kernel_busy_ms"),
                    "hbm_gbs_over_kernel_time": prof["hbm_bytes_per_assembly"] / (prof["kernel_busy_ms"] * 1e-3) / 1e9
                    if prof.get("kernel_busy_ms") else None,
                    "traffic_source": prof["source"]})
    except (OSError, ValueError, KeyError):
        pass
    return rec


def ensemble_run(steps):
    """The 256-realization batched ensemble (N log-uniform in [64, 2048],
    cutoff multiplier 1, precision 1e-6, zero stress), autotuned, one
    graph-captured attempt per replay."""
    from sdddst_amd import configs
    from sdddst_amd.ensemble import Ensemble
    sizes, x, y, b = configs.ensemble_workload(256, 3)
    with Ensemble(sizes, x, y, b, cutoff_multiplier=1.0, precision=1e-6, field_mode="fast") as e:
        report = e.autotune()
        e.run(1)
        t0 = time.perf_counter()
        attempts = e.run(steps + 1)
        wall = time.perf_counter() - t0
        t, h, s, f = e.state()
        st = e.stats()
    done = int(s.sum()) - len(sizes)
    return {"workload": "256 realizations, N log-uniform in [64, 2048] (seed 3), m=1, precision 1e-6, zero stress",
            "steps_per_realization": steps, "sum_n": int(sum(sizes)), "wall_s": wall, "attempts": attempts,
            "ms_per_attempt": wall / attempts * 1e3, "realization_steps_per_s": done / wall,
            "failed_attempts": int(f.sum()), "solver_iterations_per_solve": st.solver_iterations / max(1, st.solves),
            "lu_fallbacks": st.lu_fallbacks,
            "autotuned": {"speeds": "tiles" if st.speeds_block == 0 else f"per-target block {st.speeds_block}",
                          "solver_block": st.solver_block, "spmv_lanes_per_row": st.spmv_lanes},
            "autotune_report": report.strip().split("\n")}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def timed_steps(sim, k, label):
    """Wall time of k successful steps (failed attempts included)."""
    s0 = sim.state()
    ok = 0
    t0 = time.perf_counter()
    last = t0
    while ok < k:
        ok += 1 if sim.step() else 0
        now = time.perf_counter()
        if now - last > 10:
            log(f"{label}: {ok}/{k} successful steps, {now - t0:.1f} s")
            last = now
    t1 = time.perf_counter()
    s1 = sim.state()
    return t1 - t0, s1.succesful_steps - s0.succesful_steps, s1.failed_steps - s0.failed_steps


def host_cpu():
    """Model, thread count and C library of the host the baseline runs on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count()
    libc, ver = platform.libc_ver()
    return {"model": model, "nproc": os.cpu_count(), "threads_available": threads, "libc": f"{libc} {ver}"}


def cpu_baseline(x, y, b, seconds):
    """The oracle (the CPU restatement of the reference, 1 thread, -O2, no
    contraction) on bounded samples, on this host:
    - the small configuration in full: N = 256 (seed 1), m = 1, zero stress,
      precision 1e-6, 100 successful steps, dense LU standing in for UMFPACK
      (the reference's own CPU-runnable case);
    - the headline workload (N = 16384, m = 2, constant stress 0.02): the
      pair terms of a row sample at full N, scaled to the 5 speed evaluations
      of a successful step, plus one full Jacobian assembly timed, times the
      3 assemblies of a step.  The sparse factorisations and solves
      (UMFPACK, absent from the host) are not timed, so `value` is an upper
      bound on the reference's step rate there."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import ctypes

    import oracle_lib as O
    from sdddst_amd import configs
    lib = O.oracle()
    # small configuration, in full
    x1, y1, b1 = configs.dislocations(256, 1)
    t0 = time.perf_counter()
    s1 = O.OracleSim(x1, y1, b1, cutoff_multiplier=1.0, prec=1e-6, h0=1e-6, stress="zero")
    while s1.succesful_steps < 100:
        s1.step()
    c1_wall = time.perf_counter() - t0
    c1 = {"steps_per_s": 100 / c1_wall, "wall_s": c1_wall, "failed_steps": int(s1.failed_steps),
          "sim_time": s1.time}
    s1.close()
    # headline workload: pair-term row sample
    n = len(x)
    pairs_total = n * (n - 1) / 2
    rows = 16
    done_pairs = 0
    elapsed = 0.0
    i0 = 0
    while elapsed < seconds:
        i1 = min(n, i0 + rows)
        cnt = ctypes.c_long()
        t0 = time.perf_counter()
        lib.orc_speeds_rows_sample(n, O.dp(x), O.dp(y), O.dp(b), i0, i1, ctypes.byref(cnt))
        elapsed += time.perf_counter() - t0
        done_pairs += cnt.value
        # spread the sample over the triangle, wrapping around: the stride
        # visits every 16-row block in turn
        i0 = (i1 + n // 64) % n
    pairs_per_s = done_pairs / elapsed
    # one full assembly of the headline matrix
    t0 = time.perf_counter()
    Ap = O.jacobian(x, y, b, 2.0, 1e-6, cap=n * 700)[0]
    asm_s = time.perf_counter() - t0
    step_s = 5 * pairs_total / pairs_per_s + 3 * asm_s
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"headline workload: oracle pair terms of {done_pairs} pairs (row sample of N={n}) in "
                      f"{elapsed:.1f} s = {pairs_per_s:.3e} pairs/s, x5 evaluations per step, plus one full "
                      f"oracle Jacobian assembly ({int(Ap[-1])} entries) in {asm_s:.1f} s, x3 per step; "
                      "factorisation and solves not timed (no UMFPACK on the host); and the N=256 configuration "
                      f"run in full ({c1_wall:.1f} s)",
            "pairs_per_s": pairs_per_s, "assembly_s": asm_s, "speeds_eval_s": pairs_total / pairs_per_s,
            "config1_n256_full_run": c1, "cpu": host_cpu()}


def ev_cpu_baseline(x, y, b, seconds_cap=60.0):
    """The CPU side of one EV snapshot as the reference runs it, minus the
    eigenvalue-derivative loop: the Hessian from the oracle's tau_x (numpy
    diagonal) and LAPACK's symmetric eigensolver through numpy, on the
    threads numpy's BLAS uses.  The reference's derivative loop is O(N^4)
    (N^4 tensor terms per snapshot) and is not timed."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    n = len(x)
    t0 = time.perf_counter()
    I, J = np.tril_indices(n, -1)
    d, e = x[I] - x[J], y[I] - y[J]
    _, xyd = O.xy_batch(d - np.floor(d + 0.5), e - np.floor(e + 0.5))
    H = np.zeros((n, n))
    H[I, J] = b[I] * b[J] * xyd
    H[J, I] = H[I, J]
    np.fill_diagonal(H, -H.sum(axis=1))
    t1 = time.perf_counter()
    np.linalg.eigh(H)
    t2 = time.perf_counter()
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count()
    return {"hessian_s": t1 - t0, "eigh_s": t2 - t1, "kind": "port", "cores": threads,
            "sample": f"oracle tau_x Hessian of the same N={n} snapshot + numpy.linalg.eigh (LAPACK syevd); "
                      "derivative loop (O(N^4)) not timed"}


def ev_run(args):
    """One EV snapshot on the device: sdddst_ev_decompose (dense Hessian,
    rocSOLVER dsyevd, per-mode observables) and sdddst_ev_analysis with the
    eigenvalue derivatives; one untimed call first (library set-up), then
    the mean of three."""
    import sdddst_amd as S
    from sdddst_amd import configs
    n = args.ev_analysis
    x, y, b = configs.dislocations(n, 2)
    with S.Context(n) as c:
        c.upload_state(0, x, y, b)
        v = c.speeds(0, S.V_SCRATCH, sigma=0.0, update_tol=False)
        c.ev_decompose(0, speeds=v, stress_rate=1.0)
        times = {}
        for label, kw in (("decompose_s", {}), ("with_derivatives_s", dict(speeds=v, stress_rate=1.0))):
            t0 = time.perf_counter()
            for _ in range(3):
                out = c.ev_decompose(0, **kw)
            times[label] = (time.perf_counter() - t0) / 3
    rec = {"ev_analysis": {"n": n, "data": "synthetic (SplitMix64 seed 2 positions, balanced +/-1 Burgers vectors)",
                           **times, "eigenvalue_range": [float(out["eigenvalues"][0]), float(out["eigenvalues"][-1])],
                           "reference_derivative_terms": float(n) ** 4}}
    if not args.no_cpu_baseline:
        rec["ev_analysis"]["cpu_baseline"] = ev_cpu_baseline(x, y, b)
    print(json.dumps(rec), flush=True)


def large_run(args):
    """N = 65536 dislocations (seed 5) + 2048 point defects (seed 5) with
    A and KASQR scaled as the command line scales them, m = 2, precision
    1e-6, fixed-rate stress 4e6, learned preconditioner; successful steps
    per second of a run from t = 0, and the pair kernel's rate."""
    from sdddst_amd import configs
    from sdddst_amd._capi import kernel_stats, timing
    from sdddst_amd.simulation import Simulation
    n, pc, rate = 65536, 2048, 4.0e6
    steps = args.steps if args.steps != 200 else 100
    x, y, b = configs.dislocations(n, 5)
    px, py = configs.point_defects(n, pc, 5)
    A = (1e-4 * 16.0) * (1.0 / np.sqrt(n))
    K = 1.65 * 1.65 * 1e6 / 256.0 * n
    log(f"workload n={n} point defects={pc} fixed-rate {rate:g}")
    sim = Simulation(x, y, b, cutoff_multiplier=2.0, precision=1e-6, initial_step=1e-6, stress="fixed-rate",
                     stress_value=rate, solver="fgmres", preconditioner=args.precond, px=px, py=py, A=A, KASQR=K,
                     solver_tol=1e-10, solver_max_iter=300, solver_restart=40)
    ctx = sim.device_context
    timing(ctx, True, pair_only=True)
    wall, ok, failed = timed_steps(sim, steps, "timed")
    pair_cnt, pair_ms = kernel_stats(ctx, "pair")
    timing(ctx, False)
    st = sim.state()
    sim.close()
    pairs = n * (n - 1) / 2
    avg = pair_ms / max(pair_cnt, 1)
    line = {"metric": "successful implicit steps/s at N=65536 + 2048 point defects, 1e-6 precision",
            "value": ok / wall, "unit": "steps/s", "n_gpus": 1, "steps": ok, "warmup": 0,
            "ms_per_step": wall / ok * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (SplitMix64 seed 5 positions and point defects)",
            "config": {"workload": "n65536-pd2048-m2-ramp", "n": n, "point_defects": pc, "cutoff_multiplier": 2.0,
                       "precision": 1e-6, "stress": f"fixed-rate {rate:g}", "initial_step": 1e-6,
                       "preconditioner": args.precond, "failed_steps_in_timed_region": failed,
                       "final_time": st.sim_time, "final_stress": rate * st.sim_time},
            "roofline": dict(pair_roofline(pairs, avg, None, None), launches=pair_cnt,
                             launch_sampling="HIP events on one pair-kernel launch in 4 of the timed run"),
            "solver_iterations_per_solve": st.solver_total_iters / max(1, st.solver_total_solves),
            "solver_fallbacks_total": st.solver_fallbacks,
            "solver_precond_fallbacks_total": st.solver_precond_fallbacks,
            "solver_strong_retries_total": st.solver_strong_retries,
            "attempts_rejected_for_unconverged_solve": st.aborted_attempts}
    print(json.dumps(line))


def m4_run(args):
    """N = 16384 (seed 4), cutoff multiplier 4 (about 50 within-cutoff
    neighbours per row), zero stress, precision 1e-6: the learned
    preconditioner against the pair blocks alone and against the direct
    solver, each from t = 0."""
    from sdddst_amd import configs
    n = 16384
    x, y, b = configs.dislocations(n, 4)
    steps = args.steps
    out = {}
    for label, solver, precond, k in (("learned", "fgmres", "learned", steps), ("pair", "fgmres", "pair", steps),
                                      ("direct", "dense-lu", "pair", max(1, args.plain_steps))):
        from sdddst_amd.simulation import Simulation
        sim = Simulation(x, y, b, cutoff_multiplier=4.0, precision=1e-6, initial_step=1e-6, stress="zero",
                         solver=solver, preconditioner=precond, solver_tol=1e-10, solver_max_iter=300,
                         solver_restart=40)
        if args.warmup > 0 and label != "direct":
            timed_steps(sim, 1, f"{label} warmup")
            sim.close()
            sim = Simulation(x, y, b, cutoff_multiplier=4.0, precision=1e-6, initial_step=1e-6, stress="zero",
                             solver=solver, preconditioner=precond, solver_tol=1e-10, solver_max_iter=300,
                             solver_restart=40)
        s0 = sim.state()
        w, ok, failed = timed_steps(sim, k, label)
        st = sim.state()
        sv = st.solver_total_solves - s0.solver_total_solves
        out[label] = {"steps_per_s": ok / w, "steps": ok, "failed": failed,
                      "solver_iterations_per_solve": (st.solver_total_iters - s0.solver_total_iters) / sv if sv
                      else None, "precond_fallbacks": st.solver_precond_fallbacks,
                      "direct_fallbacks": st.solver_fallbacks}
        sim.close()
        log(f"{label}: {ok / w:.2f} steps/s")
    line = {"metric": "successful implicit steps/s at N=16384, cutoff multiplier 4, 1e-6 precision",
            "value": out["learned"]["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": out["learned"]["steps"],
            "warmup": args.warmup, "ms_per_step": 1e3 / out["learned"]["steps_per_s"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (SplitMix64 seed 4 positions, balanced +/-1 Burgers vectors)",
            "config": {"workload": "n16384-m4-zero-learned", "n": n, "cutoff_multiplier": 4.0, "precision": 1e-6,
                       "stress": 0.0, "initial_step": 1e-6},
            "learned": out["learned"], "pair_block_preconditioner": out["pair"], "plain_direct_solver": out["direct"],
            "learned_over_pair": out["learned"]["steps_per_s"] / out["pair"]["steps_per_s"],
            "learned_over_direct": out["learned"]["steps_per_s"] / out["direct"]["steps_per_s"]}
    print(json.dumps(line))


def main():
    args = parse()
    if args.trace_gaps:
        trace_gaps(args.trace_gaps)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.assembly_only:
        print(json.dumps({"assembly": assembly_run(args.n, 2.0, max(1, args.assembly_reps))}))
        return
    if args.ev_analysis > 0:
        if args.gpus != 1:
            print("bench.py: this build measures one GPU", file=sys.stderr)
            sys.exit(2)
        ev_run(args)
        return
    if args.workload == "n16384-m4-zero-learned":
        if args.gpus != 1:
            print("bench.py: this build measures one GPU", file=sys.stderr)
            sys.exit(2)
        m4_run(args)
        return
    if args.workload == "n65536-pd2048-m2-ramp":
        if args.gpus != 1:
            print("bench.py: this build measures one GPU", file=sys.stderr)
            sys.exit(2)
        large_run(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(one process drives one MI355X; see the sharding design)", file=sys.stderr)
        sys.exit(2)
    from sdddst_amd import configs
    from sdddst_amd._capi import kernel_stats, timing
    n = args.n
    x, y, b = configs.dislocations(n, 2)

    log(f"workload n={n} solver={args.solver} field={args.field_mode}")
    # Warm-up steps run on a separate simulation of the same workload (device
    # code paths, library handles, allocations); the timed run then starts
    # from the initial configuration at t = 0 with its inputs already in HBM,
    # as the workload's run of `steps` successful steps does.
    # The timed run records only the pair kernel's events (every event pair
    # idles the stream briefly); the whole-speeds and assembly averages come
    # from the warm-up run.
    sp_cnt = sp_ms = jac_cnt = jac_ms = 0
    if args.warmup > 0:
        wsim = make_sim(args, args.solver, x, y, b)
        timing(wsim.device_context, True)
        timed_steps(wsim, args.warmup, "warmup")
        sp_cnt, sp_ms = kernel_stats(wsim.device_context, "speeds")
        jac_cnt, jac_ms = kernel_stats(wsim.device_context, "jacobian")
        timing(wsim.device_context, False)
        wsim.close()
    sim = make_sim(args, args.solver, x, y, b)
    ctx = sim.device_context
    timing(ctx, True, pair_only=True)
    s_before = sim.state()
    wall, ok, failed = timed_steps(sim, args.steps, "timed")
    log(f"timed: {ok} steps in {wall:.2f} s")
    pair_cnt, pair_ms = kernel_stats(ctx, "pair")
    timing(ctx, False)
    st = sim.state()
    solver_name = sim.solver_name
    solves = st.solver_total_solves - s_before.solver_total_solves
    iters = st.solver_total_iters - s_before.solver_total_iters
    sim.close()

    steps_per_s = ok / wall
    traffic, traffic_rec = traffic_record()
    pairs = n * (n - 1) / 2
    avg_pair_ms = pair_ms / max(pair_cnt, 1)
    line = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": ok,
        "warmup": args.warmup,
        "ms_per_step": wall / ok * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 2 positions, balanced +/-1 Burgers vectors)",
        "config": {"workload": "n16384-m2-const0.02", "n": n, "cutoff_multiplier": 2.0, "precision": 1e-6,
                   "stress": 0.02, "initial_step": 1e-6, "solver": solver_name, "preconditioner": args.precond,
                   "field_mode": args.field_mode,
                   "failed_steps_in_timed_region": failed},
        "roofline": dict(pair_roofline(pairs, avg_pair_ms, traffic, traffic_rec), launches=pair_cnt,
                         launch_sampling="HIP events on one pair-kernel launch in 4 of the timed run"),
        # every pair of the workload over the whole speeds evaluation: the
        # series of the x-separated tiles, the pair kernel on the others,
        # the preparation and the finishing pass
        "pairs_per_s": pairs / (sp_ms / max(sp_cnt, 1) * 1e-3) if sp_cnt else None,
        "speeds_eval_ms": sp_ms / sp_cnt if sp_cnt else None,
        "jacobian_ms": jac_ms / jac_cnt if jac_cnt else None,
        "final_step_size": st.step_size,
        "solver_iterations_per_solve": iters / solves if solves else None,
        "solver_fallbacks_total": st.solver_fallbacks,
        "solver_precond_fallbacks_total": st.solver_precond_fallbacks,
    }
    line["roofline"]["issue_probe"] = issue_probe(line["roofline"])
    if args.assembly_reps > 0:
        log("Jacobian assembly")
        line["assembly"] = assembly_run(n, 2.0, args.assembly_reps)
    if args.compare_steps > 0 and args.solver == "fgmres" and args.precond != "pair":
        log("pair-block preconditioner (the method without the network)")
        sim3 = make_sim(args, "fgmres", x, y, b, precond="pair")
        s3 = sim3.state()
        w3, ok3, f3 = timed_steps(sim3, args.compare_steps, "pair timed")
        st3 = sim3.state()
        sv = st3.solver_total_solves - s3.solver_total_solves
        line["pair_block_preconditioner"] = {
            "steps_per_s": ok3 / w3, "steps": ok3, "failed": f3,
            "solver_iterations_per_solve": (st3.solver_total_iters - s3.solver_total_iters) / sv if sv else None,
            "headline_speedup": steps_per_s / (ok3 / w3)}
        sim3.close()
    if args.plain_steps > 0 and args.solver != "dense-lu":
        log("direct-solver path")
        sim2 = make_sim(args, "dense-lu", x, y, b)
        timed_steps(sim2, 1, "plain warmup")
        w2, ok2, f2 = timed_steps(sim2, args.plain_steps, "plain timed")
        line["plain_direct_solver"] = {"solver": sim2.solver_name, "steps_per_s": ok2 / w2, "steps": ok2,
                                       "failed": f2,
                                       "speedup_of_headline": steps_per_s / (ok2 / w2)}
        sim2.close()
    if args.ensemble_steps > 0:
        log("batched ensemble")
        line["ensemble_256"] = ensemble_run(args.ensemble_steps)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(x, y, b, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
