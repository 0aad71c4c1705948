#!/usr/bin/env python3
"""Benchmark of the PeriDyn explicit-dynamics hot path on one MI355X.

Default workload (the largest single-GPU configuration): 256 independent
tensile specimens (bond-based and elastic ordinary state-based alternating,
2k-60k particles each, log-uniform from seed 20261203; 4.4M particles and 4.6e8
bonds in all), stepped together in the FP32-compensated form: one pd_run per
call, one launch per kernel class per step, steady steps replayed from a
captured HIP graph.  Every specimen's grips move from step 1, so every block's
deformed latch is set on step 1 and every timed step runs the bond passes.  A
"step" is one velocity_verlet_step! over all 256 specimens; all state is
resident in HBM before timing starts.  By default the timed run is each
workload's own run length (500 steps here, after 5 untimed ones); every bond of
the families counts as one bond-update per step, broken or not, as the
reference's bond loop visits every bond and tests its mask.

Reported on one JSON line:
  value          bond-updates/s of the timed (graph-replayed) steps
  roofline       the dominant kernel's algorithmic bytes per launch over its
                 mean launch time, from HIP events around every launch on the
                 library stream in an eager run of the same steps
  cpu_baseline   the oracle (C++ restatement, threaded loops) on a sample of
                 the same specimens on the host cores
  fp64_one_at_a_time   the same specimens run one at a time in plain FP64
  served         the served stream (FP32C continuous batching by default)
                 against the same requests served one at a time in plain FP64

Other workloads (--workload):
  plates_ep    two EP plates 50x50x20 (100k particles), nonlinear contact
  disk_impact  BB disk r=150 t=10 + stiff sphere (721k particles), linear contact
  tensile_bb   BB tensile block 20^3 (8k particles), MoveBC grips
  served       the served stream alone (sims/s, p50/p99 time-to-result)
The specimens of plates_ep and disk_impact move rigidly until they touch
(their deformed latches skip the bond passes), so those workloads first step
untimed past first contact (--advance) and time the steps after it.
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
# timed steps per workload by default: the run length of its configuration
RUN_STEPS = {"batched": 500, "plates_ep": 500, "disk_impact": 2000, "tensile_bb": 200}
PROF_CLASSES = ("prologue", "latch", "theta", "td", "force_bb", "force_osb", "force_ep", "contact", "book")

# Algorithmic bytes per launch, (per bond, per particle), of each kernel class
# and form: int32 column + uint8 state per bond, neighbour gathers counted as
# cache hits, edp a per-bond stream; the Verlet update (read v, f, y; write v,
# y: 120 B/particle) is fused into the force kernel.  FP32C reads a 16 B x32
# record instead of the 32 B FP64 one (and 8 B of FP32 theta, 1/m for OSB); its
# elastoplastic passes read the FP32 records and keep the FP64 plastic
# extensions (8 B read + 8 B write per bond in the force pass, 8 B read in td).
BYTES = {
    ("force_bb", "fp64"): (5, 204),
    ("force_osb", "fp64"): (5, 220),
    ("force_ep", "fp64"): (21, 228),
    ("theta", "fp64"): (5, 72),
    ("td", "fp64"): (13, 84),
    ("force_bb", "fp32c"): (5, 188),
    ("force_osb", "fp32c"): (5, 196),
    ("force_ep", "fp32c"): (21, 212),
    ("theta", "fp32c"): (5, 56),
    ("td", "fp32c"): (13, 68),
}

# untimed steps before the timed region: the rigid approach of the colliding
# workloads (plates: first contact at step 23; disk: at step 6)
ADVANCE = {"plates_ep": 30, "disk_impact": 10, "tensile_bb": 0}


def form_of(precision: str) -> str:
    return "fp32c" if precision == "fp32c" else "fp64"


def workload(name: str):
    from peridyn_amd import workloads as W
    if name == "plates_ep":
        return ("two EP plates 50x50x20 (100k particles), delta=3.015dx, nonlinear contact",
                W.plates_ep(), 10, "force_ep")
    if name == "disk_impact":
        return ("BB disk r=150dx t=10dx + sphere r=15dx (721k particles), linear contact",
                W.disk_impact(), 10, "force_bb")
    if name == "tensile_bb":
        return ("BB tensile block 20^3 (8k particles), MoveBC grips, dt=1e-3",
                W.tensile_bb(), 1 << 30, "force_bb")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(spec, nuf: int, budget_s: float, nbonds: int):
    """The oracle (C++ restatement) on the first steps of the same workload
    until ~budget_s of stepping has elapsed, with the loops the reference runs
    under Threads.@threads (force, dilatation, td_norm2, contact search) on
    OpenMP threads (OMP_NUM_THREADS, else every host core).  The deformed latch
    of every block is set first so that the sample times steady-state steps
    (bond loops running), as the GPU's timed steps after first contact do."""
    from oracle import oracle as O
    O.build()
    env = O.OracleEnv(spec)
    env.threaded = True
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
    cores = env.threads_used
    return {"value": nbonds * steps / el, "unit": "bond-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {steps} steps of the same workload with every block's deformed latch "
                      f"set ({el:.1f} s), C++ restatement of the reference loops (oracle/) with its threaded "
                      f"loops on {cores} OpenMP threads, not Julia"}


def cpu_baseline_batched(specs, steps: int, budget_s: float):
    """The oracle on an evenly strided sample of the batch, each sampled
    specimen for the same number of steps as the GPU run, until ~budget_s of
    stepping has elapsed (family builds are not timed)."""
    from oracle import oracle as O
    O.build()
    order = list(range(0, len(specs), 8)) + [i for i in range(len(specs)) if i % 8]
    done, bsteps, el, cores = 0, 0, 0.0, 1
    parts = 0
    for i in order:
        env = O.OracleEnv(specs[i])
        env.threaded = True
        nb = int(sum(int(b.intact.sum()) for b in env.blocks))
        parts += env.N
        env.update_neighs()
        env.apply_bc_at0()
        t0 = time.perf_counter()
        for _ in range(steps):
            env.step()
        el += time.perf_counter() - t0
        cores = env.threads_used
        bsteps += nb * steps
        done += 1
        if el >= budget_s:
            break
    return {"value": bsteps / el, "unit": "bond-updates/s", "cores": cores, "kind": "port",
            "sample": f"{done} of the {len(specs)} specimens ({parts} particles), {steps} steps each "
                      f"({el:.1f} s of stepping), C++ restatement of the reference loops (oracle/) with its "
                      f"threaded loops on {cores} OpenMP threads, not Julia"}


def traffic_for(kernel: str, workload_name: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t[workload_name][kernel]["hbm_bytes_per_launch"]
    except Exception:
        return None


def roofline(prof: dict, dom: str, form: str, nnz: int, N: int, workload_name: str):
    """Roofline record of the dominant kernel class `dom` from the per-class
    HIP-event totals of the timed region."""
    ms_dom, n_dom = prof[dom]
    bb, bp = BYTES[(dom, form)]
    alg = bb * nnz + bp * N
    avg_s = (ms_dom / max(n_dom, 1)) / 1e3
    achieved = alg / avg_s / 1e9 if avg_s > 0 else 0.0
    kern = {"force_bb": "BB", "force_osb": "OSB", "force_ep": "EP"}.get(dom, dom)
    return {"bound": "hbm", "kernel": f"{dom} ({kern}, {form})", "achieved": achieved, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_for(dom, workload_name),
            "alg_bytes_per_launch": alg, "bytes_per_bond": bb, "bytes_per_particle": bp,
            "avg_launch_ms": avg_s * 1e3, "launches": n_dom}


def check_line(out: dict, steps: int):
    """A line whose figures cannot be real is an error, not a result: the
    dominant kernel's HBM fraction must be at most 1, and its time per step at
    most the measured time per step."""
    rf = out["roofline"]
    per_step = rf["avg_launch_ms"] * rf["launches"] / max(steps, 1)
    problems = []
    if not (0.0 < rf["frac"] <= 1.0):
        problems.append(f"roofline frac {rf['frac']:.3f} outside (0, 1]")
    if per_step > out["eager_profiled_run"]["ms_per_step"] * 1.05:
        problems.append(f"dominant kernel {per_step:.3f} ms per step exceeds the step time")
    if problems:
        print(json.dumps({"error": "; ".join(problems), "line": out}), file=sys.stderr)
        raise SystemExit("bench line rejected: " + "; ".join(problems))


def read_profile(ctx):
    prof = {k: ctx.profile_read(k) for k in PROF_CLASSES}
    ctx.profile(False)
    return prof


def copy_calibration():
    """A plain 256 MiB device copy (x3) whose HBM counter readings scale
    FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc."""
    import torch
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def bench_single(args):
    import torch
    from peridyn_amd import core

    desc, spec, nuf, dom = workload(args.workload)
    prec = args.precision or "fp64"
    adv = ADVANCE[args.workload] if args.advance < 0 else args.advance
    if args.copy_calibration:
        copy_calibration()
    ctx = core.Context(0)
    lanes = args.lanes
    tuned = None
    if args.autotune and prec != "exact":
        from peridyn_amd import autotune
        tuned = autotune.tune(spec, prec, steps=10, nuf=nuf, use_cache=False, ctx=ctx)
        lanes = tuned["lanes_per_particle"]
    # warm-up on a throwaway copy (code objects, caches)
    if args.warmup > 0:
        w = core.Env.from_spec(spec, precision=prec, ctx=ctx, lanes_per_particle=lanes)
        w.run(args.warmup, nuf)
        ctx.sync()
        w.close()

    def advanced_env():
        e = core.Env.from_spec(spec, precision=prec, ctx=ctx, lanes_per_particle=lanes)
        if adv > 0:
            e.run(adv, nuf)
        ctx.sync()
        return e

    # timed region: the run as a caller gets it (steady steps replayed from the
    # captured HIP graph, no per-kernel events), after the untimed approach
    env = advanced_env()
    lanes = env.lanes_per_particle
    info = env.info()
    N, nnz = info["particles"], info["bonds"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    env.run(args.steps, nuf)
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    env.check_finite()
    env.close()
    # kernel durations: the same steps launched eagerly with HIP events around
    # every launch on the library stream (graph replays carry no per-kernel
    # events; a kernel's own duration does not depend on how it was launched)
    env = advanced_env()
    ctx.profile(True)
    t0 = time.perf_counter()
    env.run(args.steps, nuf)
    ctx.sync()
    el_prof = time.perf_counter() - t0
    prof = read_profile(ctx)
    env.close()
    logged = None
    if args.log_every > 0:
        env = advanced_env()
        t0 = time.perf_counter()
        rows = env.run_logged(args.steps, nuf, args.log_every)
        ctx.sync()
        el_log = time.perf_counter() - t0
        env.close()
        logged = {"every": args.log_every, "rows": int(rows.shape[0]), "ms_per_step": 1e3 * el_log / args.steps,
                  "last_row": [float(v) for v in rows[-1]]}
    form = form_of(prec)
    value = nnz * args.steps / el
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{desc}; {args.steps} timed steps after {adv} untimed steps of approach",
                   "particles": N, "bonds": nnz, "precision": prec, "neigh_update_freq": nuf,
                   "lanes_per_particle": lanes, "device_bytes": info["device_bytes"],
                   "autotune_ms_per_step": tuned["ms_per_step"] if tuned else None},
        "roofline": dict(roofline(prof, dom, form, nnz, N, args.workload),
                         measured_in="eager run of the same steps, HIP events around every launch"),
        "kernel_ms": {k: round(v[0], 3) for k, v in prof.items() if v[1] > 0},
        "eager_profiled_run": {"value": nnz * args.steps / el_prof, "ms_per_step": 1e3 * el_prof / args.steps},
    }
    if logged:
        out["logged_run"] = logged
    check_line(out, args.steps)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(spec, nuf, args.cpu_budget, nnz)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    if args.served_requests > 0:
        out["served"] = served_summary(ctx, args.served_requests, args.served_rate, seq=True,
                                       precision=args.served_precision, lanes=args.served_lanes)
    print(json.dumps(out))


def bench_batched(args):
    """The batch of specimens stepped together (one launch per kernel class
    per step), FP32C by default, against the same specimens one at a time in
    plain FP64."""
    import torch
    from peridyn_amd import core
    from peridyn_amd import workloads as W

    if args.copy_calibration:
        copy_calibration()
    specs = W.batched_specimens(args.batch_count)
    ctx = core.Context(0)
    prec = args.precision or "fp32c"
    if prec == "exact":
        raise SystemExit("the batched workload runs the fp64 or fp32c form")
    # warm-up: code objects and allocator on a few specimens
    w = [core.Env.from_spec(s, precision=prec, ctx=ctx) for s in specs[:4]]
    core.run_batch(w, max(args.warmup, 1))
    for e in w:
        e.close()
    lanes = args.lanes
    tuned = None
    if lanes == 0:
        from peridyn_amd import autotune
        tuned = autotune.tune_batch(specs, prec, candidates=(2, 4, 8), steps=5, sample=32, ctx=ctx)
        lanes = tuned["lanes_per_particle"]
        print(f"batch schedule: {tuned}", file=sys.stderr, flush=True)

    def build():
        es = [core.Env.from_spec(s, precision=prec, ctx=ctx, lanes_per_particle=lanes) for s in specs]
        ctx.sync()
        return es

    # timed region: the run as a caller gets it (steady steps replayed from the
    # captured HIP graph)
    envs = build()
    infos = [e.info() for e in envs]
    N = sum(i["particles"] for i in infos)
    nnz = sum(i["bonds"] for i in infos)
    dev_bytes = sum(i["device_bytes"] for i in infos)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    core.run_batch(envs, args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t0
    for e in envs:
        e.check_finite()
        e.close()
    print(f"batched graph run: {1e3 * el_g / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    # the same steps launched eagerly with HIP events around every launch
    envs = build()
    ctx.profile(True)
    t0 = time.perf_counter()
    core.run_batch(envs, args.steps)
    ctx.sync()
    el = time.perf_counter() - t0
    prof = read_profile(ctx)
    for e in envs:
        e.close()
    dom = max(("force_bb", "force_osb", "force_ep"), key=lambda k: prof[k][0])
    # nnz/N of the kernel class's own blocks
    kind = {"force_bb": "bb", "force_osb": "osb", "force_ep": "ep"}[dom]
    kn = sum(i["bonds"] for s, i in zip(specs, infos) if s["blocks"][0]["kind"] == kind)
    kp = sum(i["particles"] for s, i in zip(specs, infos) if s["blocks"][0]["kind"] == kind)
    # plain FP64, one specimen at a time (each its own pd_run)
    seq_t = 0.0
    seq_nnz = 0
    ns = len(specs) if args.seq_sample <= 0 else min(args.seq_sample, len(specs))
    stride = max(1, len(specs) // ns)
    sample = specs[::stride][:ns]
    for s in sample:
        e = core.Env.from_spec(s, precision="fp64", ctx=ctx)
        seq_nnz += e.info()["bonds"]
        ctx.sync()
        t0 = time.perf_counter()
        e.run(args.steps)
        ctx.sync()
        seq_t += time.perf_counter() - t0
        e.close()
    print(f"fp64 one at a time: {seq_nnz * args.steps / seq_t:.3e} bond-updates/s", file=sys.stderr, flush=True)
    seq_value = seq_nnz * args.steps / seq_t
    value = nnz * args.steps / el_g
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el_g / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} tensile specimens (BB/OSB 50/50, 2k-60k particles, seed 20261203), "
                               f"{prec}, one pd_run over all, {args.steps} timed steps from step 1",
                   "particles": N, "bonds": nnz,
                   "precision": prec, "device_bytes": dev_bytes, "lanes_per_particle": lanes,
                   "batch_tune_ms_per_step": tuned["ms_per_step"] if tuned else None},
        "eager_profiled_run": {"value": nnz * args.steps / el, "ms_per_step": 1e3 * el / args.steps},
        "roofline": dict(roofline(prof, dom, form_of(prec), kn, kp, "batched"),
                         measured_in="eager run of the same steps, HIP events around every launch"),
        "kernel_ms": {k: round(v[0], 3) for k, v in prof.items() if v[1] > 0},
        "fp64_one_at_a_time": {"value": seq_value, "specimens_timed": len(sample),
                               "note": "evenly strided sample of the same specimens, same steps, plain FP64, "
                                       "one pd_run per specimen"},
        "vs_fp64_one_at_a_time": value / seq_value,
    }
    check_line(out, args.steps)
    if not args.no_cpu_baseline:
        cb = cpu_baseline_batched(specs, args.steps, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    if args.served_requests > 0:
        out["served"] = served_summary(ctx, args.served_requests, args.served_rate, seq=True,
                                       precision=args.served_precision, lanes=args.served_lanes)
    print(json.dumps(out))


def _progress(label: str, total: int):
    def cb(res):
        if res.id % 16 == 15:
            print(f"{label}: request {res.id} done at {res.done:.2f} s (latency {res.latency:.3f} s)",
                  file=sys.stderr, flush=True)
    return cb


def served_summary(ctx, count: int, rate: float, seq: bool = True, seed: int = 20261204, lanes: int = 0,
                   precision: str = "fp64"):
    """Serve `count` requests (continuous batching in `precision`) and, if
    `seq`, the same stream one request at a time in plain FP64.  `lanes`: lanes
    per particle of the served envs (0: each env's own choice).  One untimed
    pass of the whole stream first leaves the context's device-memory cache as
    a long-running server has it, for both timed passes."""
    from peridyn_amd import serve
    from peridyn_amd import workloads as W

    reqs = W.served_requests(count=count, seed=seed, rate=rate)
    for r in reqs:
        r["spec"] = W.make_spec(r)
    free, total = ctx.mem_info()
    budget = int(0.8 * free)
    # warm-up: code objects and the device-memory cache
    serve.Server(serve.GpuBackend(ctx, precision, keep_state=False, lanes_per_particle=lanes), budget,
                 chunk=50).serve(reqs)
    srv = serve.Server(serve.GpuBackend(ctx, precision, keep_state=False, lanes_per_particle=lanes), budget,
                       chunk=50, on_retire=_progress(f"served {precision}", count))
    _, st = srv.serve(reqs)
    out = {"requests": count, "arrival_rate_per_s": rate, "precision": precision, "chunk_steps": 50,
           "lanes_per_particle": lanes,
           "budget_bytes": budget, **st.as_dict()}
    if seq:
        one = serve.Server(serve.GpuBackend(ctx, "fp64", keep_state=False), budget, chunk=None, max_active=1,
                           on_retire=_progress("fp64 one at a time", count), overlap=False)
        _, s1 = one.serve(reqs)
        out["fp64_one_at_a_time"] = s1.as_dict()
        out["sims_per_s_vs_fp64_one_at_a_time"] = st.sims_per_s / s1.sims_per_s
        out["p99_vs_fp64_one_at_a_time"] = st.p99 / s1.p99
    return out


def bench_served(args):
    from peridyn_amd import core

    ctx = core.Context(0)
    s = served_summary(ctx, args.served_requests or 64, args.served_rate, seq=not args.no_seq,
                       lanes=args.served_lanes, precision=args.served_precision)
    out = {
        "metric": METRIC,
        "value": s["sims_per_s"],
        "unit": "sims/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 2,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.served_precision == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{s['requests']} Poisson requests at {args.served_rate}/s: tensile / indentation "
                               "50/50, BB / OSB 50/50, 1k-200k particles, 200-2000 steps, horizon 2.5-4.0 dx; "
                               f"{args.served_precision.upper()} continuous batching, 50-step chunks"},
        "served": s,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def make_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's own run length: batched 500, plates_ep 500, "
                         "disk_impact 2000, tensile_bb 200)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "plates_ep", "disk_impact", "tensile_bb", "served"])
    ap.add_argument("--precision", default=None, choices=["fp64", "exact", "fp32c"],
                    help="form of the timed run (default: fp32c for batched, fp64 otherwise)")
    ap.add_argument("--advance", type=int, default=-1,
                    help="plates_ep / disk_impact / tensile_bb: untimed steps before the timed ones "
                         "(-1: past first contact)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per particle (0: chosen at upload / by the tuner)")
    ap.add_argument("--autotune", action="store_true", help="time every schedule first and use the fastest")
    ap.add_argument("--log-every", type=int, default=0,
                    help="also time the run logging the printdata columns every K steps")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--seq-sample", type=int, default=64,
                    help="batched: specimens timed one at a time in FP64 (0: all)")
    ap.add_argument("--served-requests", type=int, default=None,
                    help="requests of the served stream (default 64 for batched and served, 0 otherwise; "
                         "0 skips the served summary)")
    ap.add_argument("--served-rate", type=float, default=200.0,
                    help="served: Poisson arrivals per second (the default exceeds the GPU's service rate, so "
                         "requests queue and the GPU stays saturated)")
    ap.add_argument("--served-lanes", type=int, default=4,
                    help="served: lanes per particle of the served envs (0: each env's own choice)")
    ap.add_argument("--served-precision", default="fp32c", choices=["fp64", "fp32c"],
                    help="served: form of the batched simulations")
    ap.add_argument("--no-seq", action="store_true", help="served: skip the one-at-a-time FP64 baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (x3) whose HBM counter readings "
                         "scale FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc")
    return ap


def main():
    args = make_parser().parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this build has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = RUN_STEPS.get(args.workload, 100)
    if args.served_requests is None:
        args.served_requests = 64 if args.workload in ("batched", "served") else 0
    if args.workload == "batched":
        bench_batched(args)
    elif args.workload == "served":
        bench_served(args)
    else:
        bench_single(args)


if __name__ == "__main__":
    main()
