#!/usr/bin/env python3
"""Benchmark of the PeriDyn explicit-dynamics hot path on one MI355X.

Workload (default): two elastoplastic (ordinary state-based, Von Mises) plates of
50x50x20 lattice particles (100,000 particles, 10,975,216 bonds) approaching at
0.05 dx/step with a nonlinear spring contact (cutoff 0.9 dx, lists rebuilt every
10 steps), 500 velocity-Verlet steps.  A "step" is one velocity_verlet_step!
over the whole specimen pair; all state is resident in HBM before timing starts.

Reported: bond-updates/s (initial bonds x steps / time), the HBM-roofline
fraction of the dominant kernel (EP force pass, timed with HIP events on the
library stream inside the timed region), and the CPU oracle timed on the host
on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "bond-updates/s & % HBM roofline; served sims/s + p99 latency vs plain FP64"

# algorithmic bytes per launch, per bond / per particle, of each kernel class
# (FP64 forms; int32 column + uint8 state per bond, neighbour gathers counted
# as cache hits, edp a per-bond stream; Verlet traffic folded into the fused
# force kernel).
BYTES = {
    ("force_bb", "fp64"): (5, 204),
    ("force_osb", "fp64"): (5, 220),
    ("force_ep", "fp64"): (21, 228),
    ("theta", "fp64"): (5, 72),
    ("td", "fp64"): (13, 84),
}


def workload(name: str):
    from peridyn_amd import workloads as W
    if name == "plates_ep":
        return ("two EP plates 50x50x20 (100k particles), delta=3.015dx, nonlinear contact, 500 steps",
                W.plates_ep(), 10, "force_ep")
    if name == "disk_impact":
        return ("BB disk r=150dx t=10dx + sphere r=15dx (721k particles), linear contact",
                W.disk_impact(), 10, "force_bb")
    if name == "tensile_bb":
        return ("BB tensile block 20^3 (8k particles), MoveBC grips, dt=1e-3",
                W.tensile_bb(), 1 << 30, "force_bb")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(spec, nuf: int, budget_s: float, nbonds: int):
    """The oracle (C++ restatement, single thread) on the first steps of the
    same workload until ~budget_s of stepping has elapsed.  The deformed latch
    of every block is set first so that the sample times steady-state steps
    (bond loops running), as in almost all of the GPU's timed steps; the first
    steps of the workload, with rigidly moving blocks, would skip them."""
    from oracle import oracle as O
    O.build()
    env = O.OracleEnv(spec)
    env.update_neighs()
    env.apply_bc_at0()
    for b in env.blocks:
        b.deformed = True
    steps = 0
    t0 = time.perf_counter()
    while True:
        env.step()
        steps += 1
        if steps % nuf == 0:
            env.update_neighs()
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 500:
            break
    return {"value": nbonds * steps / el, "unit": "bond-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {steps} steps of the same workload with every block's deformed latch "
                      f"set ({el:.1f} s), single-threaded C++ restatement of the reference loops (oracle/), "
                      "not Julia"}


def traffic_for(kernel: str, workload_name: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        e = t[workload_name][kernel]
        return e["hbm_bytes_per_launch"]
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="plates_ep")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "exact", "fp32c"])
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per particle (0: chosen at upload)")
    ap.add_argument("--autotune", action="store_true", help="time every schedule first and use the fastest")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (x3) whose HBM counter readings "
                         "scale FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, "
              "'How the path would shard'); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from peridyn_amd import core

    desc, spec, nuf, dom = workload(args.workload)
    if args.copy_calibration:
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ctx = core.Context(0)
    lanes = args.lanes
    tuned = None
    if args.autotune and args.precision != "exact":
        from peridyn_amd import autotune
        tuned = autotune.tune(spec, args.precision, steps=10, nuf=nuf, use_cache=False, ctx=ctx)
        lanes = tuned["lanes_per_particle"]
    # warm-up on a throwaway copy (code objects, caches), then time the canonical run
    if args.warmup > 0:
        w = core.Env.from_spec(spec, precision=args.precision, ctx=ctx, lanes_per_particle=lanes)
        w.run(args.warmup, nuf)
        ctx.sync()
        w.close()
    env = core.Env.from_spec(spec, precision=args.precision, ctx=ctx, lanes_per_particle=lanes)
    lanes = env.lanes_per_particle
    info = env.info()
    N, nnz = info["particles"], info["bonds"]
    ctx.sync()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    env.run(args.steps, nuf)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    el = t1 - t0
    prof = {k: ctx.profile_read(k) for k in ("prologue", "latch", "theta", "td", "force_bb", "force_osb",
                                             "force_ep", "contact", "book")}
    ctx.profile(False)
    env.check_finite()

    value = nnz * args.steps / el
    ms_step = 1e3 * el / args.steps
    ms_dom, n_dom = prof[dom]
    bb, bp = BYTES[(dom, "fp64")]
    alg_bytes = bb * nnz + bp * N
    avg_s = (ms_dom / max(n_dom, 1)) / 1e3
    achieved = alg_bytes / avg_s / 1e9 if avg_s > 0 else 0.0
    # whole-step algorithmic bytes (all passes) against the step time
    step_bytes = sum(BYTES[(k, "fp64")][0] * nnz + BYTES[(k, "fp64")][1] * N
                     for k in ("theta", "td", "force_ep") if prof[k][1] > 0) if dom == "force_ep" else alg_bytes
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "particles": N, "bonds": nnz, "precision": args.precision,
                   "neigh_update_freq": nuf, "lanes_per_particle": lanes,
                   "autotune_ms_per_step": tuned["ms_per_step"] if tuned else None},
        "roofline": {"bound": "hbm", "kernel": f"k_force<{dom[6:].upper()},{args.precision}>",
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_for(dom, args.workload),
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": avg_s * 1e3,
                     "launches": n_dom},
        "step_hbm_frac": (step_bytes / (el / args.steps)) / 1e9 / PEAK_HBM_GBS,
        "kernel_ms": {k: round(v[0], 3) for k, v in prof.items() if v[1] > 0},
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(spec, nuf, args.cpu_budget, nnz)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
