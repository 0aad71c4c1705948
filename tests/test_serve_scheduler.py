"""Serving loop scheduling (host logic, no GPU): admission by footprint,
chunking at multiples of the neighbour-update period, retirement, latency
accounting, with a simulated clock and a stand-in backend."""
import math

import numpy as np
import pytest

from peridyn_amd import serve
from peridyn_amd import workloads as W


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


class FakeBackend:
    """Each step costs `per_step` seconds per chunk (batched: independent of how
    many simulations share it); admit costs `per_admit`."""

    def __init__(self, clock, per_step=1e-3, per_admit=1e-3, free=1 << 40):
        self.clock = clock
        self.per_step = per_step
        self.per_admit = per_admit
        self.free = free
        self.resident = {}
        self.steps_done = {}
        self.calls = []
        self.next = 0

    def admit(self, spec):
        h = self.next
        self.next += 1
        nb = serve.footprint_estimate(spec) // 2
        self.resident[h] = nb
        self.steps_done[h] = 0
        self.clock.t += self.per_admit
        return h, nb, 100 * spec["blocks"][0]["x"].shape[0]

    def step(self, handles, n, nuf):
        self.calls.append((tuple(handles), n, nuf))
        for h in handles:
            self.steps_done[h] += n
        self.clock.t += self.per_step * n

    def retire(self, h, result):
        result.y = np.array([self.steps_done[h]])
        del self.resident[h]

    def free_bytes(self):
        return self.free - sum(self.resident.values())


def small_requests(count=12, rate=200.0, seed=5):
    reqs = W.served_requests(count=count, seed=seed, nmin=300, nmax=3000, smin=20, smax=200, rate=rate)
    for r in reqs:
        r["spec"] = W.make_spec(r)
    return reqs


def test_footprint_estimate_bounds_lattice_sizes():
    spec = W.tensile(8000)
    est = serve.footprint_estimate(spec)
    n = spec["blocks"][0]["x"].shape[0]
    # at least the bond stream of a full-degree lattice and the particle state
    assert est >= n * W.lattice_ball_count(3.015) * 5 + n * 200
    # EP blocks carry rev + edp per slot
    ep = W.tensile_bb(n=10, kind="ep")
    bb = W.tensile_bb(n=10, kind="bb")
    assert serve.footprint_estimate(ep) > 3 * serve.footprint_estimate(bb) // 2


def test_every_request_completes_with_its_steps_and_chunks_are_aligned():
    clock = FakeClock()
    be = FakeBackend(clock)
    srv = serve.Server(be, budget_bytes=1 << 40, chunk=50, clock=clock, sleep=clock.sleep, reserve_bytes=0,
                       keep_log=True)
    reqs = small_requests()
    res, st = srv.serve(reqs)
    assert st.completed == len(reqs)
    for r in reqs:
        assert int(res[r["id"]].y[0]) == r["steps"]
        assert res[r["id"]].done >= res[r["id"]].admitted >= r["arrival"]
    for handles, n, nuf in be.calls:
        assert n % nuf == 0 and 0 < n <= 50
    assert st.max_active > 1  # requests overlapped
    lat = sorted(r.latency for r in res.values())
    assert st.p50 == serve.percentile(lat, 50) and st.p99 == lat[-1]


def test_admission_respects_the_budget():
    clock = FakeClock()
    be = FakeBackend(clock)
    reqs = small_requests(count=16, rate=1e6)   # all arrive at once
    biggest = max(serve.footprint_estimate(r["spec"]) for r in reqs)
    budget = int(2.2 * biggest)
    srv = serve.Server(be, budget_bytes=budget, chunk=50, clock=clock, sleep=clock.sleep, reserve_bytes=0,
                       keep_log=True)
    res, st = srv.serve(reqs)
    assert st.completed == len(reqs)
    assert st.peak_bytes <= budget
    for ev in st.log:
        if ev[0] in ("admit", "retire"):
            assert ev[3] <= budget


def test_request_larger_than_budget_is_rejected():
    clock = FakeClock()
    be = FakeBackend(clock)
    reqs = small_requests(count=3)
    srv = serve.Server(be, budget_bytes=1000, clock=clock, sleep=clock.sleep, reserve_bytes=0)
    with pytest.raises(ValueError):
        srv.serve(reqs)


def test_steps_must_be_multiples_of_nuf():
    clock = FakeClock()
    reqs = small_requests(count=2)
    reqs[0]["steps"] += 3
    srv = serve.Server(FakeBackend(clock), budget_bytes=1 << 40, clock=clock, sleep=clock.sleep)
    with pytest.raises(ValueError):
        srv.serve(reqs)


def test_one_at_a_time_mode_runs_each_request_in_one_call():
    clock = FakeClock()
    be = FakeBackend(clock)
    reqs = small_requests(count=6)
    srv = serve.Server(be, budget_bytes=1 << 40, chunk=None, max_active=1, clock=clock, sleep=clock.sleep,
                       reserve_bytes=0)
    res, st = srv.serve(reqs)
    assert st.max_active == 1 and len(be.calls) == len(reqs)
    assert [c[1] for c in be.calls] == [r["steps"] for r in sorted(reqs, key=lambda r: r["arrival"])]


def test_idle_server_waits_for_arrivals():
    clock = FakeClock()
    be = FakeBackend(clock, per_step=1e-6)
    reqs = small_requests(count=4, rate=1.0)  # ~1 s apart, each served in ~ms
    srv = serve.Server(be, budget_bytes=1 << 40, clock=clock, sleep=clock.sleep, reserve_bytes=0)
    res, st = srv.serve(reqs)
    for r in reqs:
        assert res[r["id"]].admitted >= r["arrival"]
        assert res[r["id"]].latency < 0.1
    assert math.isclose(st.makespan, max(x.done for x in res.values()) - min(r["arrival"] for r in reqs))


class FakeAsyncBackend(FakeBackend):
    """FakeBackend whose chunks run in the background: step_begin returns at
    once, the chunk completes per_step * n seconds later on the clock."""

    def __init__(self, clock, **kw):
        super().__init__(clock, **kw)
        self.end_at = None
        self.chunk_spans = []   # (start, end, handles)
        self.admits = []        # (clock at admission, handle)

    def admit(self, spec):
        h, nb, bonds = super().admit(spec)
        self.admits.append((self.clock.t, h))
        return h, nb, bonds

    def step_begin(self, handles, n, nuf):
        assert self.end_at is None, "a chunk is already running"
        self.calls.append((tuple(handles), n, nuf))
        for h in handles:
            self.steps_done[h] += n
        self.end_at = self.clock.t + self.per_step * n
        self.chunk_spans.append((self.clock.t, self.end_at, tuple(handles)))

    def step_done(self):
        return self.clock.t >= self.end_at

    def step_end(self):
        self.clock.t = max(self.clock.t, self.end_at)
        self.end_at = None

    def retire(self, h, result):
        assert self.end_at is None, "retired while a chunk runs"
        super().retire(h, result)


def test_overlapped_admission_joins_the_next_chunk():
    clock = FakeClock()
    be = FakeAsyncBackend(clock, per_step=2e-4, per_admit=1e-3)
    reqs = small_requests(count=24, rate=400.0, seed=11)
    srv = serve.Server(be, budget_bytes=1 << 40, chunk=50, clock=clock, sleep=clock.sleep, reserve_bytes=0,
                       keep_log=True)
    res, st = srv.serve(reqs)
    assert st.completed == len(reqs)
    for r in reqs:
        assert int(res[r["id"]].y[0]) == r["steps"]
        assert res[r["id"]].done >= res[r["id"]].admitted >= r["arrival"]
    # some admissions happened while a chunk ran, and a simulation admitted
    # during a chunk is not stepped by it
    inside = 0
    for t, h in be.admits:
        for (a, b, hs) in be.chunk_spans:
            if a < t < b:
                inside += 1
                assert h not in hs
    assert inside > 0 and st.admit_overlap_s > 0
    for handles, n, nuf in be.calls:
        assert n % nuf == 0 and 0 < n <= 50


def test_overlap_can_be_turned_off():
    clock = FakeClock()
    be = FakeAsyncBackend(clock)
    reqs = small_requests(count=6)
    srv = serve.Server(be, budget_bytes=1 << 40, chunk=None, max_active=1, clock=clock, sleep=clock.sleep,
                       reserve_bytes=0, overlap=False)
    res, st = srv.serve(reqs)
    assert st.completed == len(reqs) and not be.chunk_spans and st.admit_overlap_s == 0
    with pytest.raises(ValueError):
        serve.Server(FakeBackend(clock), budget_bytes=1 << 40, overlap=True)
