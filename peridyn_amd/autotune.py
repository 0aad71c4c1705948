"""Schedule autotuner for the bond loops.

The bond layout (pd_internal.h) serves each particle with G lanes (G = 1 is one
lane per particle, G = 64 one wavefront per particle).  Which G is fastest
depends on the specimen size (enough wavefronts to hide gather latency), the
degree distribution and the material passes.  `tune(spec, precision)` builds a
replica of the env for each candidate G, times a few steps of the real run
loop (wall clock around pd_run, which returns when its steps are done), and
returns the fastest G together with the timings; results are cached per
(particle-count bucket, material kinds, precision) in a JSON file so later envs
of similar shape reuse them.  `tune_batch(specs, precision)` does the same for
many specimens stepped together by one pd_run, where the batch as a whole, not
each specimen, has to fill the chip.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Dict, Iterable, Optional

import numpy as np

from . import core

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)
_CACHE_PATH = os.environ.get("PERIDYN_TUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "peridyn_amd_tune.json"))


def _key(spec: dict, precision: str) -> str:
    n = sum(int(b["x"].shape[0]) for b in spec["blocks"])
    kinds = "+".join(sorted({b["kind"] for b in spec["blocks"]}))
    bucket = int(round(math.log2(max(n, 1)) * 2))  # half-octave buckets
    return f"{precision}:{kinds}:{bucket}"


def _load() -> Dict[str, dict]:
    try:
        with open(_CACHE_PATH) as f:
            return json.load(f)
    except Exception:
        return {}


def _store(db: Dict[str, dict]):
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        with open(_CACHE_PATH, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass


def _deformed_replica(spec: dict) -> dict:
    """Copy of spec whose blocks are deformed from the first step (the deformed
    latch would otherwise skip the bond passes of a rigidly moving specimen)."""
    rep = dict(spec)
    rep["blocks"] = []
    for b in spec["blocks"]:
        b2 = dict(b)
        y = np.array(b.get("y", b["x"]), dtype=np.float64, copy=True)
        y[:, 0] += 1e-7 * np.sin(np.arange(y.shape[0]))
        b2["y"] = y
        rep["blocks"].append(b2)
    return rep


def time_schedule(spec: dict, precision: str, lanes: int, steps: int = 20, nuf: int = 10,
                  ctx: Optional[core.Context] = None) -> float:
    """Mean wall time per step (ms) of `steps` steps with the given schedule."""
    ctx = ctx or core.Context.default()
    env = core.Env.from_spec(_deformed_replica(spec), precision=precision, ctx=ctx, lanes_per_particle=lanes)
    try:
        env.run(2, nuf)  # warm code objects and caches
        ctx.sync()
        t0 = time.perf_counter()
        env.run(steps, nuf)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / steps
    finally:
        env.close()


def tune(spec: dict, precision: str = "fp64", candidates: Iterable[int] = CANDIDATES, steps: int = 20,
         nuf: int = 10, use_cache: bool = True, ctx: Optional[core.Context] = None) -> dict:
    """Returns {"lanes_per_particle": best G, "ms_per_step": {G: ms}}."""
    if precision == "exact":
        return {"lanes_per_particle": 1, "ms_per_step": {}}
    key = _key(spec, precision)
    db = _load() if use_cache else {}
    if use_cache and key in db:
        return db[key]
    times = {int(g): time_schedule(spec, precision, int(g), steps, nuf, ctx) for g in candidates}
    best = min(times, key=times.get)
    res = {"lanes_per_particle": best, "ms_per_step": times}
    if use_cache:
        db[key] = res
        _store(db)
    return res


def time_batch(specs: list, precision: str, lanes: int, steps: int = 10, nuf: int = 10,
               ctx: Optional[core.Context] = None) -> float:
    """Mean wall time per step (ms) of the specimens stepped together."""
    ctx = ctx or core.Context.default()
    envs = []
    try:
        for s in specs:
            envs.append(core.Env.from_spec(_deformed_replica(s), precision=precision, ctx=ctx,
                                           lanes_per_particle=lanes))
        core.run_batch(envs, 2, nuf)
        ctx.sync()
        t0 = time.perf_counter()
        core.run_batch(envs, steps, nuf)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / steps
    finally:
        for e in envs:
            e.close()


def tune_batch(specs: list, precision: str = "fp32c", candidates: Iterable[int] = (1, 2, 4, 8, 16),
               steps: int = 10, nuf: int = 10, sample: int = 128, ctx: Optional[core.Context] = None) -> dict:
    """Lanes per particle for a batch: every candidate G timed on an evenly
    strided sample of the specimens.  Returns {"lanes_per_particle", "ms_per_step",
    "sample"}."""
    if precision == "exact":
        return {"lanes_per_particle": 1, "ms_per_step": {}, "sample": 0}
    stride = max(1, len(specs) // max(sample, 1))
    sub = specs[::stride][:sample]
    times = {int(g): time_batch(sub, precision, int(g), steps, nuf, ctx) for g in candidates}
    best = min(times, key=times.get)
    return {"lanes_per_particle": best, "ms_per_step": times, "sample": len(sub)}
