"""Continuous-batching request loop over device-resident simulations.

A request is one simulation: a spec (or the arguments of
peridyn_amd.workloads.make_spec), a step count and a contact-rebuild period
`nuf`.  The loop:

  * admits arrived requests in arrival order while their device footprint fits
    the budget: an upper-bound estimate before upload (footprint_estimate),
    replaced by the library's own count (pd_env_info) once the env is resident;
  * steps every resident simulation together, one pd_run over all of them per
    chunk (one launch per kernel class covers every specimen);
  * retires each simulation the moment its steps are done (results copied to
    the host, device memory freed) and admits waiting ones into the room;
  * with a backend that can start a chunk and return (step_begin / step_done /
    step_end), admits the requests that arrive while a chunk runs during that
    chunk: their upload and family build overlap the stepping on the device,
    and they join the next chunk.

Chunks are multiples of `nuf` and never pass a resident request's last step, so
each request rebuilds contacts at the same steps as when it runs alone in one
run! call (run! counts steps from 1 in each call and rebuilds at
i % neigh_update_freq == 0, src/solvers/solvers.jl:133,177).  Specimens share
no data arrays.  Tests compare served results with the same request run alone:
they are bit-identical, for every request of the full benchmark stream.

The scheduling logic takes a backend object, so it runs unchanged over the HIP
library (GpuBackend) or a host-side stand-in in tests.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

# per-particle device bytes of an env (x, y[2], v, f, aux, x32/aux32 as double4 /
# float4, plus scalars), per bond slot (col + state; EP adds rev + edp[2])
_BYTES_PARTICLE = 7 * 32 + 64
_BYTES_SLOT = 5
_BYTES_SLOT_EP = 5 + 4 + 16


def footprint_estimate(spec: dict) -> int:
    """Upper bound of the device bytes pd_env_commit will hold for `spec`.

    Bond slots are bounded per block by (n + 63) * (dmax + 64): every wavefront
    chunk is padded to its widest row, rounded up to the lane-group size.  dmax
    is the unit-lattice ball count when the block is an integer lattice (its
    exact maximum degree) and max_neigh - 1 otherwise (larger degrees are an
    overflow error).  Contacts add the own lists (n1 * max_neighs) and the
    transposed lists."""
    from .workloads import lattice_ball_count
    total = 0
    n_all = 0
    for b in spec["blocks"]:
        x = np.asarray(b["x"])
        n = x.shape[0]
        n_all += n
        if np.array_equal(x, np.round(x)) and np.all(np.asarray(b["vol"]) == 1.0):
            dmax = lattice_ball_count(float(b["horizon"]))
        else:
            dmax = int(b.get("max_neigh", 200)) - 1
        per_slot = _BYTES_SLOT_EP if b["kind"] == "ep" else _BYTES_SLOT
        total += (n + 63) * (dmax + 64) * per_slot
    total += n_all * _BYTES_PARTICLE
    for c in spec.get("contacts", []):
        total += 3 * n_all * 4 + 2 * n_all * int(c.get("max_neighs", 50)) * 4 + 4096
    return int(total * 1.1) + (1 << 20)


@dataclass
class Result:
    id: int
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    steps: int = 0
    bonds: int = 0
    y: Optional[np.ndarray] = None
    v: Optional[np.ndarray] = None
    f: Optional[np.ndarray] = None
    damage: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Active:
    req: dict
    handle: object
    bytes: int
    left: int
    result: Result


@dataclass
class Stats:
    completed: int = 0
    makespan: float = 0.0
    sims_per_s: float = 0.0
    bond_updates_per_s: float = 0.0
    p50: float = 0.0
    p99: float = 0.0
    mean_active: float = 0.0
    max_active: int = 0
    peak_bytes: int = 0
    chunks: int = 0
    admit_s: float = 0.0    # host + device time spent admitting (upload, family build)
    admit_overlap_s: float = 0.0  # the part of admit_s spent while a chunk was running
    step_s: float = 0.0     # time inside pd_run
    retire_s: float = 0.0   # result download and free
    idle_s: float = 0.0     # waiting for arrivals with nothing resident
    log: List[tuple] = field(default_factory=list)

    def as_dict(self) -> dict:
        return {"completed": self.completed, "makespan_s": self.makespan, "sims_per_s": self.sims_per_s,
                "bond_updates_per_s": self.bond_updates_per_s, "p50_s": self.p50, "p99_s": self.p99,
                "mean_active": self.mean_active, "max_active": self.max_active, "peak_bytes": self.peak_bytes,
                "chunks": self.chunks, "admit_s": self.admit_s, "admit_overlap_s": self.admit_overlap_s,
                "step_s": self.step_s, "retire_s": self.retire_s, "idle_s": self.idle_s}


def percentile(a, q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    a = sorted(a)
    if not a:
        return math.nan
    k = max(1, int(math.ceil(q / 100.0 * len(a))))
    return float(a[k - 1])


class Server:
    """Request loop.  `backend` provides

        admit(req) -> (handle, device_bytes, bonds)   upload and build
        step(handles, nsteps, nuf)                     one pd_run over them all
        retire(handle, result)                         fill result, free memory
        free_bytes() -> int                            device memory now free

    and, to overlap admission with stepping, optionally

        step_begin(handles, nsteps, nuf)   start the chunk and return
        step_done() -> bool                whether it has completed
        step_end()                         wait for it and complete it

    `chunk` is the largest number of steps per pd_run (None: run each batch to
    the first completion); `max_active` caps the resident simulations;
    `poll_s` is how often arrivals are checked while a chunk runs; `overlap`
    (default: whenever the backend can) admits during chunks."""

    def __init__(self, backend, budget_bytes: int, chunk: Optional[int] = 50, max_active: int = 1 << 30,
                 clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                 reserve_bytes: int = 1 << 30, keep_log: bool = False,
                 on_retire: Optional[Callable[[Result], None]] = None, poll_s: float = 2e-4,
                 overlap: Optional[bool] = None):
        self.backend = backend
        self.budget = int(budget_bytes)
        self.chunk = chunk
        self.max_active = max_active
        self.clock = clock
        self.sleep = sleep
        self.reserve = reserve_bytes
        self.keep_log = keep_log
        self.on_retire = on_retire
        self.poll_s = poll_s
        can = all(hasattr(backend, m) for m in ("step_begin", "step_done", "step_end"))
        if overlap and not can:
            raise ValueError("the backend cannot start a chunk and return")
        self.overlap = can if overlap is None else bool(overlap)

    def serve(self, requests: List[dict], make_spec: Optional[Callable[[dict], dict]] = None):
        """Serve `requests` (dicts with id, arrival [s after start], steps, nuf,
        and spec or make_spec's arguments).  Returns ({id: Result}, Stats)."""
        nufs = {int(r["nuf"]) for r in requests}
        if len(nufs) > 1:
            raise ValueError("requests stepped together must share one neighbour-update period")
        nuf = nufs.pop() if nufs else 1
        for r in requests:
            if int(r["steps"]) <= 0 or int(r["steps"]) % nuf:
                raise ValueError(f"request {r['id']}: steps must be a positive multiple of nuf={nuf}")
        if self.chunk is not None and self.chunk % nuf:
            raise ValueError("chunk must be a multiple of nuf")
        pending = sorted(requests, key=lambda r: (r["arrival"], r["id"]))
        specs = {}
        ests = {}
        for r in pending:
            spec = r["spec"] if "spec" in r else make_spec(r)
            specs[r["id"]] = spec
            ests[r["id"]] = footprint_estimate(spec)
            if ests[r["id"]] > self.budget:
                raise ValueError(f"request {r['id']} needs ~{ests[r['id']]} device bytes, more than the "
                                 f"budget {self.budget}")
        active: List[_Active] = []
        results: Dict[int, Result] = {}
        st = Stats()
        used = 0
        t0 = self.clock()
        busy_t = 0.0
        active_t = 0.0
        running = None  # overlapped chunk in flight: (stepped simulations, k, start time)

        def admit_arrived(now: float):
            nonlocal used
            i = 0
            while i < len(pending) and len(active) < self.max_active:
                r = pending[i]
                if r["arrival"] > now:
                    break
                est = ests[r["id"]]
                if used + est > self.budget or est > self.backend.free_bytes() - self.reserve:
                    i += 1
                    continue
                pending.pop(i)
                res = Result(id=r["id"], arrival=r["arrival"], steps=int(r["steps"]))
                ta = self.clock()
                handle, nbytes, bonds = self.backend.admit(specs[r["id"]])
                res.admitted = self.clock() - t0
                st.admit_s += res.admitted - (ta - t0)
                if running is not None:
                    st.admit_overlap_s += res.admitted - (ta - t0)
                res.bonds = bonds
                used += max(nbytes, 0)
                active.append(_Active(r, handle, nbytes, int(r["steps"]), res))
                st.peak_bytes = max(st.peak_bytes, used)
                if self.keep_log:
                    st.log.append(("admit", r["id"], res.admitted, used))

        def chunk_done(stepped: List[_Active], k: int, dt: float):
            nonlocal used, busy_t, active_t, active
            busy_t += dt
            st.step_s += dt
            active_t += dt * len(stepped)
            st.chunks += 1
            st.max_active = max(st.max_active, len(stepped))
            if self.keep_log:
                st.log.append(("step", tuple(a.req["id"] for a in stepped), k))
            finished = set()
            for a in stepped:
                a.left -= k
                if a.left == 0:
                    tr = self.clock()
                    self.backend.retire(a.handle, a.result)
                    a.result.done = self.clock() - t0
                    st.retire_s += a.result.done - (tr - t0)
                    used -= max(a.bytes, 0)
                    results[a.req["id"]] = a.result
                    finished.add(id(a))
                    if self.on_retire is not None:
                        self.on_retire(a.result)
                    if self.keep_log:
                        st.log.append(("retire", a.req["id"], a.result.done, used))
            active = [a for a in active if id(a) not in finished]

        while pending or active or running is not None:
            admit_arrived(self.clock() - t0)
            if running is not None:
                stepped, k, ts = running
                if not self.backend.step_done():
                    # every arrived request that fits was just admitted
                    now = self.clock() - t0
                    future = [r["arrival"] - now for r in pending if r["arrival"] > now]
                    self.sleep(min([self.poll_s] + future))
                    continue
                self.backend.step_end()
                running = None
                chunk_done(stepped, k, self.clock() - ts)
                continue
            if not active:
                if not pending:
                    break
                now = self.clock() - t0
                future = [r["arrival"] for r in pending if r["arrival"] > now]
                if not future:
                    raise RuntimeError("no waiting request fits the free device memory")
                wait = min(future) - (self.clock() - t0)
                if wait > 0:
                    self.sleep(wait)
                    st.idle_s += wait
                continue
            k = min(a.left for a in active)
            if self.chunk is not None:
                k = min(k, self.chunk)
            stepped = list(active)
            ts = self.clock()
            if self.overlap:
                self.backend.step_begin([a.handle for a in stepped], k, nuf)
                running = (stepped, k, ts)
                continue
            self.backend.step([a.handle for a in stepped], k, nuf)
            chunk_done(stepped, k, self.clock() - ts)
        lat = [r.latency for r in results.values()]
        st.completed = len(results)
        first = min((r.arrival for r in results.values()), default=0.0)
        last = max((r.done for r in results.values()), default=0.0)
        st.makespan = last - first
        st.sims_per_s = st.completed / st.makespan if st.makespan > 0 else math.nan
        st.bond_updates_per_s = (sum(r.bonds * r.steps for r in results.values()) / st.makespan
                                 if st.makespan > 0 else math.nan)
        st.p50 = percentile(lat, 50)
        st.p99 = percentile(lat, 99)
        st.mean_active = active_t / busy_t if busy_t > 0 else 0.0
        return results, st


class GpuBackend:
    """The HIP library as a serving backend: one env per request, all envs of a
    chunk stepped by one pd_run."""

    def __init__(self, ctx, precision: str = "fp64", keep_state: bool = True, lanes_per_particle: int = 0):
        self.ctx = ctx
        self.precision = precision
        self.keep_state = keep_state
        self.lanes = lanes_per_particle  # 0: the library's per-env choice

    def admit(self, spec: dict):
        from .core import Env
        env = Env.from_spec(spec, precision=self.precision, ctx=self.ctx, lanes_per_particle=self.lanes)
        info = env.info()
        return env, info["device_bytes"], info["bonds"]

    def step(self, handles, nsteps: int, nuf: int):
        from .core import run_batch
        run_batch(handles, nsteps, nuf)

    def step_begin(self, handles, nsteps: int, nuf: int):
        from .core import run_batch_begin
        run_batch_begin(handles, nsteps, nuf)

    def step_done(self) -> bool:
        from .core import run_batch_done
        return run_batch_done(self.ctx)

    def step_end(self):
        from .core import run_batch_end
        run_batch_end(self.ctx)

    def retire(self, env, result: Result):
        if self.keep_state:
            result.y, result.v, result.f = env.download()
            result.damage = env.damage()
        env.close()

    def free_bytes(self) -> int:
        return self.ctx.mem_info()[0]
