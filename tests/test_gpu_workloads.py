"""The benchmark workloads at full size on the GPU: family sizes against the
committed lattice fixture, the tensile block of the bond-based config
(20^3, dt = 1e-3, 200 steps) against the oracle, conservation of total
momentum by the full-size contact workloads, the FP32-compensated form
against FP64 on the largest batched specimens, and the tiled FP64 form
against the reference-order form on the full-size plates and disk, and
specimens of the full batch against the same specimens run alone.

Bars: bond counts exact; reference-order form bit-identical to the oracle over
the whole literal run; FP64 form: the literal run (dt = 1e-3 is close to the
stability limit, so rounding differences grow) is reported as the first step
past 1e-10 relative L2, and the stable twin (dt = 1e-4) stays within 1e-10."""
import json
import os
import sys

import numpy as np
import pytest

from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lattice_counts.json")))


@pytest.fixture(scope="module")
def core():
    from peridyn_amd import core
    return core


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def test_workload_bond_counts(core):
    g = core.Env.from_spec(W.tensile_bb(), precision="fp64")
    assert g.info()["bonds"] == GOLD["tensile_box_20"]["nnz"]
    g.close()
    g = core.Env.from_spec(W.plates_ep(), precision="fp64")
    assert g.info()["bonds"] == 2 * GOLD["plate_50x50x20"]["nnz"]
    g.close()
    g = core.Env.from_spec(W.disk_impact(), precision="fp64")
    assert g.info()["bonds"] == GOLD["disk_r150_t10"]["nnz"] + GOLD["sphere_r15"]["nnz"]
    g.close()


def test_tensile_literal_config_exact(core, orc):
    spec = W.tensile_bb()
    say("oracle: tensile 20^3, 200 steps")
    o = orc.OracleEnv(spec)
    o.run(200, 1 << 30)
    g = core.Env.from_spec(spec, precision="exact")
    g.run(200, 1 << 30)
    y, v, f = g.download()
    assert np.array_equal(y, o.y) and np.array_equal(v, o.v) and np.array_equal(f, o.f)
    assert np.array_equal(g.damage(), o.damage())


def test_tensile_fp64_divergence_step_and_stable_twin(core):
    spec = W.tensile_bb()
    x = spec["blocks"][0]["x"]
    e = core.Env.from_spec(spec, precision="exact")
    g = core.Env.from_spec(spec, precision="fp64")
    first = None
    for st in range(10, 201, 10):
        e.run(10, 1 << 30)
        g.run(10, 1 << 30)
        err = rel_l2(g.download()[0] - x, e.download()[0] - x)
        if first is None and err > 1e-10:
            first = st
    say(f"literal tensile config (dt = 1e-3): FP64 form first past 1e-10 at step {first} "
        f"(final rel L2 {err:.2e})")
    twin = W.tensile_bb(dt=1e-4)
    e = core.Env.from_spec(twin, precision="exact")
    g = core.Env.from_spec(twin, precision="fp64")
    e.run(200, 1 << 30)
    g.run(200, 1 << 30)
    err = rel_l2(g.download()[0] - x, e.download()[0] - x)
    say(f"stable twin (dt = 1e-4): FP64 rel L2 {err:.2e} after 200 steps")
    assert err <= 1e-10


@pytest.mark.parametrize("name", ["plates_ep", "disk_impact"])
def test_full_size_momentum_is_conserved(core, name):
    """The two contact workloads at full size (100k and 721k particles) have no
    boundary conditions, and every bond and contact force enters its two
    particles with opposite signs, so total momentum stays at its initial value
    up to rounding through collision, yielding and bond breaking.  Bar: drift
    <= 1e-9 of sum |m v| after 300 steps (FP64 form, tiled, rebuild every 10)."""
    spec = getattr(W, name)()
    m = np.concatenate([np.asarray(b["vol"], dtype=np.float64) * float(np.atleast_1d(b["density"])[0])
                        for b in spec["blocks"]])
    g = core.Env.from_spec(spec, precision="fp64")
    say(f"{name}: {g.info()['particles']} particles, layout {g.layout()}")
    _, v0, _ = g.download()
    p0 = (m[:, None] * v0).sum(axis=0)
    core.run_batch([g], 300, 10)
    y, v, f = g.download()
    assert np.all(np.isfinite(y)) and np.all(np.isfinite(v))
    p = (m[:, None] * v).sum(axis=0)
    scale = float((m[:, None] * np.abs(v)).sum())
    drift = float(np.abs(p - p0).max()) / scale
    say(f"{name}: momentum drift {drift:.2e} of sum |m v|, damage {float(g.damage().mean()):.3e}")
    assert drift <= 1e-9
    assert not np.allclose(v, v0)  # the specimens did collide
    g.close()


def test_batched_fp32c_matches_fp64_at_full_size(core):
    """The FP32-compensated form on the largest specimens of the 256-specimen
    batch (BB and OSB, up to 60k particles), batched together for the bench's
    500 steps, against the same specimens batched in plain FP64.  Bars: relative
    L2 displacement <= 1e-5 per specimen and identical bond masks."""
    specs = W.batched_specimens()
    size = [len(s["blocks"][0]["x"]) for s in specs]
    idx = sorted(range(len(specs)), key=lambda i: -size[i])[:8]
    e32 = [core.Env.from_spec(specs[i], precision="fp32c") for i in idx]
    e64 = [core.Env.from_spec(specs[i], precision="fp64") for i in idx]
    say("fp32c vs fp64, specimens", [(specs[i]["blocks"][0]["kind"], size[i]) for i in idx])
    core.run_batch(e32, 500)
    core.run_batch(e64, 500)
    worst = 0.0
    broken = 0
    for a, b, i in zip(e32, e64, idx):
        x = specs[i]["blocks"][0]["x"]
        ya, _, _ = a.download()
        yb, _, _ = b.download()
        ed = rel_l2(ya - x, yb - x)
        worst = max(worst, ed)
        _, ia = a.family(0)
        _, ib = b.family(0)
        assert np.array_equal(ia, ib), f"specimen {i}: bond masks differ"
        fam, _ = b.family(0)
        broken += int(((fam > 0) & (ib == 0)).sum())
        assert ed <= 1e-5, f"specimen {i}: displacement rel L2 {ed:.2e}"
    say(f"worst displacement rel L2 {worst:.2e}, {broken} bonds broken in all")
    for e in e32 + e64:
        e.close()


def test_plates_fp64_matches_reference_order_at_full_size(core):
    """The full-size EP plates (100k particles, nonlinear contact): the tiled FP64
    form against the reference-order form (bit-identical to the oracle on every
    smaller case) for 50 steps, through the collision and five contact
    rebuilds.  Bars: relative L2 of displacement and force <= 1e-10, identical
    bond masks.  (Measured: 1.3e-14 and 4.2e-11.  By step 60 the two forms
    differ at the contact layers only, after the rebuild at step 50: the lists
    are a discrete function of the positions, the max_neighs largest ids within
    the search distance, so a pair at that distance can enter one form's list
    and not the other's.)"""
    spec = W.plates_ep()
    ge = core.Env.from_spec(spec, precision="exact")
    gf = core.Env.from_spec(spec, precision="fp64")
    core.run_batch([ge], 50, 10)
    core.run_batch([gf], 50, 10)
    ye, ve, fe = ge.download()
    yf, vf, ff = gf.download()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    ed, ef = rel_l2(yf - x, ye - x), rel_l2(ff, fe)
    say(f"plates fp64 vs reference order: displacement rel L2 {ed:.2e}, force rel L2 {ef:.2e}")
    for k in range(2):
        _, ie = ge.family(k)
        _, i_f = gf.family(k)
        assert np.array_equal(ie, i_f), f"block {k}: bond masks differ"
    assert ed <= 1e-10 and ef <= 1e-10
    ge.close()
    gf.close()


def test_disk_fp64_matches_reference_order_at_full_size(core):
    """The full-size BB disk impact (721k particles, linear contact): the tiled
    FP64 form against the reference-order form for 50 steps with contact
    rebuilds every 10.  Bars: relative L2 of displacement and force <= 1e-10,
    identical bond masks."""
    spec = W.disk_impact()
    ge = core.Env.from_spec(spec, precision="exact")
    gf = core.Env.from_spec(spec, precision="fp64")
    core.run_batch([ge], 50, 10)
    core.run_batch([gf], 50, 10)
    ye, ve, fe = ge.download()
    yf, vf, ff = gf.download()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    ed, ef = rel_l2(yf - x, ye - x), rel_l2(ff, fe)
    say(f"disk fp64 vs reference order: displacement rel L2 {ed:.2e}, force rel L2 {ef:.2e}")
    for k in range(2):
        _, ie = ge.family(k)
        _, i_f = gf.family(k)
        assert np.array_equal(ie, i_f), f"block {k}: bond masks differ"
    assert ed <= 1e-10 and ef <= 1e-10
    ge.close()
    gf.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_full_batch_results_equal_alone(core, prec):
    """All 256 specimens of the batched workload (4.3M particles) stepped together
    for 100 steps: four of them, spread over the size range, equal the same
    specimens run alone bit for bit (specimens share no data; the lane schedule
    is the same)."""
    specs = W.batched_specimens()
    envs = [core.Env.from_spec(s, precision=prec) for s in specs]
    core.run_batch(envs, 100)
    size = [len(s["blocks"][0]["x"]) for s in specs]
    order = sorted(range(len(specs)), key=lambda i: size[i])
    picks = [order[0], order[85], order[170], order[-1]]
    say(f"{prec}: batch of {len(envs)}, checking specimens of {[size[i] for i in picks]} particles")
    for i in picks:
        alone = core.Env.from_spec(specs[i], precision=prec)
        core.run_batch([alone], 100)
        ya, va, fa = alone.download()
        yb, vb, fb = envs[i].download()
        assert np.array_equal(ya, yb) and np.array_equal(va, vb) and np.array_equal(fa, fb), f"specimen {i}"
        alone.close()
    for e in envs:
        e.close()
