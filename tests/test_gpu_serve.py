"""Served and batched forms on the GPU: every request's result is the same
when it is served in a mix (continuous batching, chunked pd_run calls over
changing sets of resident simulations) as when it runs alone in one call, and
the FP32-compensated results stay within the stated error of plain FP64.

Bars (stated here):
  * mixed vs alone, same form: bit-identical y, v, f and damage (the specimens
    share no data, chunks end on contact-rebuild steps, contact lists are
    sorted, so nothing in a result depends on its batch mates);
  * FP32C served vs FP64 alone: displacement relative L2 <= 1e-5, identical
    damage.
"""
import sys

import numpy as np
import pytest

from peridyn_amd import serve
from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def core():
    from peridyn_amd import core
    return core


def requests(count=10, seed=11):
    reqs = W.served_requests(count=count, seed=seed, nmin=400, nmax=4000, smin=30, smax=150, rate=400.0)
    for r in reqs:
        r["spec"] = W.make_spec(r)
    return reqs


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_served_mix_equals_alone(core, prec):
    reqs = requests()
    assert {r["shape"] for r in reqs} == {"tensile", "indentation"}
    ctx = core.Context(0)
    srv = serve.Server(serve.GpuBackend(ctx, precision=prec), budget_bytes=4 << 30, chunk=20)
    say(f"serve {len(reqs)} requests ({prec})")
    res, st = srv.serve(reqs)
    say(f"  done: {st.completed} sims, max resident {st.max_active}, {st.chunks} chunks")
    assert st.completed == len(reqs) and st.max_active > 1
    for r in reqs:
        e = core.Env.from_spec(r["spec"], precision=prec, ctx=ctx)
        e.run(r["steps"], r["nuf"])
        y, v, f = e.download()
        d = e.damage()
        e.close()
        got = res[r["id"]]
        assert np.array_equal(got.y, y), r["id"]
        assert np.array_equal(got.v, v), r["id"]
        assert np.array_equal(got.f, f), r["id"]
        assert np.array_equal(got.damage, d), r["id"]


def test_served_mix_equals_alone_with_poisoned_memory(core, monkeypatch):
    """The served mix in a context whose device memory is filled with 0xFF
    bytes whenever the library hands a block out (PERIDYN_POISON=1: NaN doubles,
    -1 integers), against the same requests alone in a context without it:
    bit-identical, so no kernel reads device memory it has not written."""
    reqs = requests(count=8, seed=13)
    monkeypatch.setenv("PERIDYN_POISON", "1")
    pz = core.Context(0)
    monkeypatch.delenv("PERIDYN_POISON")
    plain = core.Context(0)
    srv = serve.Server(serve.GpuBackend(pz, precision="fp64"), budget_bytes=4 << 30, chunk=20)
    res, st = srv.serve(reqs)
    assert st.completed == len(reqs) and st.max_active > 1
    for r in reqs:
        e = core.Env.from_spec(r["spec"], precision="fp64", ctx=plain)
        e.run(r["steps"], r["nuf"])
        y, v, f = e.download()
        e.close()
        got = res[r["id"]]
        assert np.array_equal(got.y, y) and np.array_equal(got.v, v) and np.array_equal(got.f, f), r["id"]
    pz.close()
    plain.close()


def test_served_fp32c_within_bar_of_fp64(core):
    reqs = requests(count=6, seed=12)
    ctx = core.Context(0)
    srv = serve.Server(serve.GpuBackend(ctx, precision="fp32c"), budget_bytes=4 << 30, chunk=20)
    res, st = srv.serve(reqs)
    worst = 0.0
    for r in reqs:
        e = core.Env.from_spec(r["spec"], precision="fp64", ctx=ctx)
        e.run(r["steps"], r["nuf"])
        y, v, f = e.download()
        d = e.damage()
        e.close()
        x = np.concatenate([b["x"] for b in r["spec"]["blocks"]])
        err = rel_l2(res[r["id"]].y - x, y - x)
        worst = max(worst, err)
        say(f"  request {r['id']} {r['shape']} {r['kind']} n={x.shape[0]}: fp32c displacement rel L2 {err:.2e}")
        assert err <= 1e-5
        assert np.array_equal(res[r["id"]].damage, d)
    say(f"worst fp32c displacement rel L2 {worst:.2e}")


def test_batched_specimens_equal_alone(core):
    """Config-3 style batch (mixed BB / OSB boxes of uneven sizes) in one
    pd_run equals each specimen alone, FP32C form."""
    specs = W.batched_specimens(count=6, seed=3, nmin=500, nmax=4000)
    ctx = core.Context(0)
    envs = [core.Env.from_spec(s, precision="fp32c", ctx=ctx) for s in specs]
    core.run_batch(envs, 40)
    for s, e in zip(specs, envs):
        a = core.Env.from_spec(s, precision="fp32c", ctx=ctx)
        a.run(40)
        for u, w in zip(e.download(), a.download()):
            assert np.array_equal(u, w)
        a.close()
        e.close()


@pytest.mark.parametrize("prec", ["fp32c", "fp64"])
def test_full_served_stream_equals_alone(core, prec):
    """The served benchmark's stream at full size (64 Poisson requests,
    1k-200k particles, 200-2000 steps, tensile and indentation, BB and OSB),
    served as the bench serves it (4 lanes per particle, 50-step chunks,
    admission overlapping the chunks; FP32C as benchmarked, and FP64): every
    request equals the same request run alone in one call, bit for bit."""
    reqs = W.served_requests(count=64, seed=20261204, rate=200.0)
    for r in reqs:
        r["spec"] = W.make_spec(r)
    ctx = core.Context(0)
    srv = serve.Server(serve.GpuBackend(ctx, prec, keep_state=True, lanes_per_particle=4),
                       int(0.8 * ctx.mem_info()[0]), chunk=50)
    say(f"serve the full stream of {len(reqs)} requests ({prec})")
    res, st = srv.serve(reqs)
    say(f"  done: {st.completed} sims, max resident {st.max_active}, {st.chunks} chunks, {st.makespan:.2f} s")
    assert st.completed == len(reqs) and st.max_active > 32
    assert st.admit_overlap_s > 0
    for r in reqs:
        e = core.Env.from_spec(r["spec"], precision=prec, ctx=ctx, lanes_per_particle=4)
        e.run(r["steps"], r["nuf"])
        y, v, f = e.download()
        d = e.damage()
        e.close()
        got = res[r["id"]]
        say(f"  request {r['id']}: {r['shape']} {r['kind']} n={r['n']} steps={r['steps']}")
        assert np.array_equal(got.y, y) and np.array_equal(got.v, v), r["id"]
        assert np.array_equal(got.f, f) and np.array_equal(got.damage, d), r["id"]


def test_latch_readback_after_a_two_block_env_with_a_pending_block(core):
    """Regression: the deformed-latch readback of a run set walks the envs in
    order and takes one flag per block.  A two-block env whose first block is
    still undeformed must consume its second block's flag too; otherwise every
    later env reads its neighbour's flags, and an env can be marked as fully
    latched while one of its blocks is not, so that block's bond forces stay
    skipped after it deforms.  Set: an indentation (both blocks rigid until the
    indenter touches), a tensile box (latches on step 1) and two plates whose
    first block is deformed from the start and whose second deforms only on
    contact (step ~23).  The plates must equal the plates run alone."""
    ctx = core.Context(0)
    a = W.indentation(2000, steps=200)
    t = W.tensile(1000)
    b = W.plates_ep(nx=8, ny=8, nz=4)
    x0 = b["blocks"][0]["x"]
    b["blocks"][0]["y"] = x0 + 1e-6 * np.sin(np.arange(x0.size)).reshape(x0.shape)
    envs = [core.Env.from_spec(s, precision="fp64", ctx=ctx) for s in (a, t, b)]
    core.run_batch(envs, 60, 10)
    alone = core.Env.from_spec(b, precision="fp64", ctx=ctx)
    alone.run(60, 10)
    ref = alone.download()
    for u, w in zip(envs[2].download(), ref):
        assert np.array_equal(u, w)
    vz2 = ref[1][x0.shape[0]:, 2]
    assert np.ptp(vz2) > 0.0  # the second plate deformed: its particles no longer move together
    for e in envs + [alone]:
        e.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_run_begin_end_equals_run_and_admits_meanwhile(core, prec):
    """pd_run_begin / pd_run_end step exactly as pd_run does (bit for bit, an
    indentation request with contacts and FixBC and a tensile one, 60 steps
    with rebuilds every 10), and an env committed while the run is in flight
    (on the admission stream) steps like one committed before it."""
    reqs = requests(count=6, seed=23)
    ind = next(r for r in reqs if r["shape"] == "indentation")
    ten = next(r for r in reqs if r["shape"] == "tensile")
    ctx = core.Context(0)
    ref = [core.Env.from_spec(r["spec"], precision=prec, ctx=ctx) for r in (ind, ten)]
    core.run_batch(ref, 60, 10)
    want = [e.download() for e in ref]
    late_ref = core.Env.from_spec(ten["spec"], precision=prec, ctx=ctx)
    late_ref.run(30, 10)
    want_late = late_ref.download()
    for e in ref + [late_ref]:
        e.close()
    say(f"async run ({prec})")
    envs = [core.Env.from_spec(r["spec"], precision=prec, ctx=ctx) for r in (ind, ten)]
    core.run_batch_begin(envs, 60, 10)
    late = core.Env.from_spec(ten["spec"], precision=prec, ctx=ctx)  # committed during the run
    core.run_batch_end(ctx)
    assert core.run_batch_done(ctx)
    for e, w in zip(envs, want):
        got = e.download()
        for a, b in zip(got, w):
            assert np.array_equal(a, b)
        assert e.time_step == 60
    late.run(30, 10)
    for a, b in zip(late.download(), want_late):
        assert np.array_equal(a, b)


def test_calls_during_an_inflight_run_are_refused(core):
    from peridyn_amd import _lib
    reqs = requests(count=4, seed=29)
    ctx = core.Context(0)
    a = core.Env.from_spec(reqs[0]["spec"], precision="fp64", ctx=ctx)
    b = core.Env.from_spec(reqs[1]["spec"], precision="fp64", ctx=ctx)
    core.run_batch_begin([a], 40, 10)
    L = ctx.lib
    assert L.pd_sync(ctx.h) == _lib.PD_E_STATE
    y = np.zeros((a.N, 3))
    assert L.pd_env_download(ctx.h, a.id, y.ctypes.data_as(_lib.P), None, None) == _lib.PD_E_STATE
    ids = np.array([b.id], dtype=np.int64)
    assert L.pd_run(ctx.h, ids.ctypes.data_as(_lib.P), 1, 10, 10) == _lib.PD_E_STATE
    assert L.pd_run_begin(ctx.h, ids.ctypes.data_as(_lib.P), 1, 10, 10) == _lib.PD_E_STATE
    assert L.pd_env_destroy(ctx.h, a.id) == _lib.PD_E_STATE  # stepped by the run
    info = b.info()  # allowed meanwhile
    assert info["particles"] == b.N
    b.close()  # not stepped by the run: allowed
    core.run_batch_end(ctx)
    assert L.pd_run_end(ctx.h) == _lib.PD_E_STATE  # nothing in flight
    a.run(10, 10)
    assert a.time_step == 50
