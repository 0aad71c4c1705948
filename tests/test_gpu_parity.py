"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars (written into each test):
  * neighbour families: identical reference ELL arrays (sorted sets, exact);
  * PD_PREC_EXACT: bit-identical y, v, f, bond state, theta, edp (== on FP64);
  * PD_PREC_FP64:  relative L2 <= 1e-10 for forces and displacements after N
    steps on stable runs; identical bond-state masks except bonds within 1e-9
    of the critical stretch.
"""
import math
import sys

import numpy as np
import pytest

from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pd():
    from peridyn_amd import core
    return core


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def run_both(pd, orc, spec, steps, nuf=1, prec="exact"):
    o = orc.OracleEnv(spec)
    o.run(steps, nuf)
    g = pd.Env.from_spec(spec, precision=prec)
    g.run(steps, nuf)
    return o, g


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["lattice9", "cloud", "disk"])
def test_family_and_weighted_volume_match_oracle(pd, orc, case):
    say("family case", case)
    if case == "lattice9":
        x, h = W.lattice_box(9, 9, 9), 3.015
    elif case == "cloud":
        x, h = np.random.default_rng(7).uniform(0.0, 9.0, size=(700, 3)), 1.7
    else:
        x, h = W.lattice_disk(12.0, 4), 3.015
    spec = {"dt": 1.0, "blocks": [W.block("bb", x, h, bulk_modulus=1.0, critical_stretch=1.0, density=1.0)],
            "contacts": [], "bcs": []}
    env = pd.Env.from_spec(spec)
    fam, intact = env.family(0)
    ofam, ointact = orc.cal_family(x, h, 200)
    assert fam.shape == ofam.shape
    assert np.array_equal(fam, ofam)
    assert np.array_equal(intact, ointact)
    m = env.weighted_volume(0)
    assert np.array_equal(m, orc.weighted_volume(x, np.ones(x.shape[0]), ofam))
    env.close()


def test_family_overflow_is_reported(pd):
    spec = {"dt": 1.0, "blocks": [W.block("bb", W.lattice_box(9, 9, 9), 3.015, max_neigh=122,
                                          bulk_modulus=1.0, critical_stretch=1.0, density=1.0)],
            "contacts": [], "bcs": []}
    from peridyn_amd._lib import PeriDynError, PD_E_OVERFLOW
    with pytest.raises(PeriDynError) as ei:
        pd.Env.from_spec(spec)
    assert ei.value.code == PD_E_OVERFLOW


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,dt,steps", [(8, 1e-4, 60), (10, 1e-3, 40)])
def test_bb_exact_bit_identical(pd, orc, n, dt, steps):
    """Tensile block with MoveBC grips; (10, 1e-3) is the literal tensile-block
    time step (unstable), where only bit-identical arithmetic can agree."""
    say("bb exact", n, dt, steps)
    spec = W.tensile_bb(n=n, dt=dt)
    o, g = run_both(pd, orc, spec, steps)
    y, v, f = g.download()
    assert np.array_equal(y, o.y)
    assert np.array_equal(v, o.v)
    assert np.array_equal(f, o.f)
    fam, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)
    assert np.array_equal(g.damage(), o.damage())
    assert g.time_step == steps


def test_bb_exact_split_runs_equal_single_run(pd, orc):
    spec = W.tensile_bb(n=7, dt=1e-4)
    o = orc.OracleEnv(spec)
    o.run(35)
    g = pd.Env.from_spec(spec, precision="exact")
    g.run(10); g.run(1); g.run(24)
    y, v, f = g.download()
    assert np.array_equal(y, o.y) and np.array_equal(v, o.v) and np.array_equal(f, o.f)


@pytest.mark.parametrize("prec", ["exact", "fp64"])
def test_tofro_reversal_on_run_boundaries(pd, orc, prec):
    """ToFroBC grips reversing every 5 steps (ToFroBC.jl:107-111), stepped in
    runs whose boundaries fall on and next to the reversal steps: equal to one
    oracle run (bit-identical in the reference-order form)."""
    spec = W.tensile_bb(n=7, dt=1e-4, pull_per_step=0.002)
    for bc in spec["bcs"]:
        bc["freq"] = 5.0
        bc["applyafter"] = 0
    o = orc.OracleEnv(spec)
    o.run(27)
    g = pd.Env.from_spec(spec, precision=prec)
    for n in (5, 1, 4, 10, 2, 5):
        g.run(n)
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    if prec == "exact":
        assert np.array_equal(y, o.y) and np.array_equal(v, o.v) and np.array_equal(f, o.f)
    else:
        assert rel_l2(y - x, o.y - x) <= 1e-10 and rel_l2(v, o.v) <= 1e-10


def test_bb_fp64_within_1e10(pd, orc):
    say("bb fp64")
    spec = W.tensile_bb(n=10, dt=1e-4, pull_per_step=0.002)
    o, g = run_both(pd, orc, spec, 100, prec="fp64")
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(f, o.f) <= 1e-10            # tolerance: 1e-10 relative L2 (FP64 form)
    assert rel_l2(y - x, o.y - x) <= 1e-10
    _, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)


@pytest.mark.parametrize("kind", ["osb", "ep"])
def test_state_based_exact_bit_identical(pd, orc, kind):
    say("state-based exact", kind)
    spec = W.tensile_bb(n=8, dt=1e-4, kind=kind, pull_per_step=0.004)
    if kind == "ep":
        spec["blocks"][0]["yield_stress"] = 20.0   # yields within the run
    o, g = run_both(pd, orc, spec, 40)
    y, v, f = g.download()
    assert np.array_equal(y, o.y)
    assert np.array_equal(v, o.v)
    assert np.array_equal(f, o.f)
    assert np.array_equal(g.theta(0), o.blocks[0].theta)
    _, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)
    if kind == "ep":
        e = g.edp(0)
        assert np.any(e != 0.0)
        assert np.array_equal(e, o.blocks[0].edp)


@pytest.mark.parametrize("kind", ["osb", "ep"])
def test_state_based_fp64_within_1e10(pd, orc, kind):
    spec = W.tensile_bb(n=8, dt=1e-4, kind=kind, pull_per_step=0.002)
    if kind == "ep":
        spec["blocks"][0]["yield_stress"] = 20.0
    o, g = run_both(pd, orc, spec, 40, prec="fp64")
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(f, o.f) <= 1e-10
    assert rel_l2(y - x, o.y - x) <= 1e-10


@pytest.mark.parametrize("expo", [1, 2, 3.0])
def test_contact_plates_exact_bit_identical(pd, orc, expo):
    """Two small EP plates colliding with a (non)linear spring contact, contact
    lists rebuilt every 5 steps.  3.0: an integer-valued Float64 exponent takes
    Julia's power by squaring, as an Int one does."""
    say("contact plates exact", expo)
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=expo)
    o, g = run_both(pd, orc, spec, 30, nuf=5)
    y, v, f = g.download()
    assert np.array_equal(y, o.y)
    assert np.array_equal(v, o.v)
    assert np.array_equal(f, o.f)
    assert np.abs(o.f).max() > 0


def test_update_acc_and_force_density_seams(pd, orc):
    spec = W.tensile_bb(n=7, kind="osb")
    rng = np.random.default_rng(2)
    spec["blocks"][0]["y"] = spec["blocks"][0]["x"] * 1.002 + rng.normal(scale=1e-3, size=(343, 3))
    o = orc.OracleEnv(spec)
    o.update_acc()
    g = pd.Env.from_spec(spec, precision="exact")
    f = g.update_acc()
    assert np.array_equal(f, o.f)
    g2 = pd.Env.from_spec(spec, precision="exact")
    f2 = np.zeros((343, 3))
    g2.force_density(0, np.ascontiguousarray(spec["blocks"][0]["y"]), f2)
    assert np.array_equal(f2, o.f)


def test_batch_equals_alone(pd, orc):
    specs = [W.tensile_bb(n=6, dt=1e-4), W.tensile_bb(n=7, dt=1e-4, kind="osb", pull_per_step=0.003)]
    alone = []
    for s in specs:
        e = pd.Env.from_spec(s, precision="exact")
        e.run(25)
        alone.append(e.download())
    envs = [pd.Env.from_spec(s, precision="exact") for s in specs]
    pd.run_batch(envs, 25)
    for e, a in zip(envs, alone):
        b = e.download()
        for u, w in zip(a, b):
            assert np.array_equal(u, w)


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
@pytest.mark.parametrize("lanes", [1, 4, 16, 64])
def test_fp64_schedules_within_1e10(pd, orc, kind, lanes):
    """Every bond-loop schedule (lanes per particle) of the FP64 form agrees with
    the oracle to 1e-10 relative L2 and keeps the exact family."""
    spec = W.tensile_bb(n=7, dt=1e-4, kind=kind, pull_per_step=0.002)
    if kind == "ep":
        spec["blocks"][0]["yield_stress"] = 20.0
    o = orc.OracleEnv(spec)
    o.run(30)
    g = pd.Env.from_spec(spec, precision="fp64", lanes_per_particle=lanes)
    assert g.lanes_per_particle == lanes
    g.run(30)
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(f, o.f) <= 1e-10
    assert rel_l2(y - x, o.y - x) <= 1e-10
    fam, intact = g.family(0)
    ofam, _ = orc.cal_family(x, 3.015, 200)
    assert np.array_equal(fam, ofam)
    assert np.array_equal(intact, o.blocks[0].intact)
    if kind == "ep":
        assert rel_l2(g.edp(0), o.blocks[0].edp) <= 1e-9


def test_contact_plates_fp64_within_1e10(pd, orc):
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    o, g = run_both(pd, orc, spec, 30, nuf=5, prec="fp64")
    y, v, f = g.download()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    assert rel_l2(f, o.f) <= 1e-10
    assert rel_l2(y - x, o.y - x) <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_contact_noninteger_exponent(pd, orc, prec):
    """Nonlinear spring with exponent 1.5 (the kernels carrying the general
    power): FP64 within the FP64 bar of the oracle (the oracle's pow is the C
    library's, the device's the ROCm math library's; both round to ~1 ulp);
    FP32C (elastoplastic plates in the FP32C form) within its bars, 1e-5 in
    displacement and 1e-4 in force."""
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    spec["contacts"][0]["exponent"] = 1.5
    o, g = run_both(pd, orc, spec, 30, nuf=5, prec=prec)
    y, v, f = g.download()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    assert np.abs(o.f).max() > 0
    bar_y, bar_f = (1e-10, 1e-10) if prec == "fp64" else (1e-5, 1e-4)
    say(f"exponent 1.5, {prec}: force {rel_l2(f, o.f):.2e}, displacement {rel_l2(y - x, o.y - x):.2e}")
    assert rel_l2(f, o.f) <= bar_f
    assert rel_l2(y - x, o.y - x) <= bar_y


def test_contact_truncation_keeps_largest_ids(pd, orc):
    """max_neighs smaller than the candidate count: the kept set is the
    max_neighs largest block-2 ids (contacts.jl:468-469), in both forms."""
    spec = W.plates_ep(nx=6, ny=6, nz=3, gap=-0.2, dt=1e-4, approach_per_step=0.0, exponent=1)
    spec["contacts"][0]["distanceX"] = 3.0
    spec["contacts"][0]["max_neighs"] = 4
    for prec in ("exact", "fp64"):
        o, g = run_both(pd, orc, spec, 3, nuf=1, prec=prec)
        y, v, f = g.download()
        if prec == "exact":
            assert np.array_equal(f, o.f)
        else:
            assert rel_l2(f, o.f) <= 1e-10


@pytest.mark.parametrize("max_neighs", [60, 2500])
def test_contact_search_large_candidate_sets(pd, orc, max_neighs):
    """Contact radius 11 dx between 24x24x12 plates: central particles have more
    than 2048 candidates (the device search's in-LDS capacity), so the kept set
    comes from the threshold search; with max_neighs = 2500 the kept lists are
    also longer than that capacity.  Exact form, bit-identical to the oracle."""
    spec = W.plates_ep(nx=24, ny=24, nz=12, gap=0.5, dt=1e-4, approach_per_step=0.0, exponent=1)
    c = spec["contacts"][0]
    c["distanceX"] = 11.0 / 0.9
    c["max_neighs"] = max_neighs
    say(f"contact search max_neighs={max_neighs}")
    o, g = run_both(pd, orc, spec, 2, nuf=1, prec="exact")
    y, v, f = g.download()
    assert np.array_equal(f, o.f)
    assert np.array_equal(y, o.y)


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
def test_fp32c_error_vs_fp64_oracle(pd, orc, kind):
    """FP32-compensated form against the FP64 oracle after 60 steps: relative
    L2 displacement error <= 1e-5, force error <= 1e-4, identical bond masks;
    for the elastoplastic block (yielding inside the run) also the plastic
    extensions within 1e-4 relative L2 (tolerances stated here; measured values
    are printed)."""
    spec = W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.003)
    if kind == "ep":
        # yielding throughout (max plastic extension 0.7) without fracture: with
        # s0 = 0.005 this block tears (26k of 60k bonds break within 60 steps)
        # and the oracle itself answers a 1e-7 perturbation of y with 1.4e-3,
        # every break decision near s0 being a discrete branch; without breaks
        # it answers linearly (1e-9 -> 4.3e-9)
        spec["blocks"][0]["yield_stress"] = 20.0
        spec["blocks"][0]["critical_stretch"] = 1.0
    o, g = run_both(pd, orc, spec, 60, prec="fp32c")
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    ed = rel_l2(y - x, o.y - x)
    ef = rel_l2(f, o.f)
    say(f"fp32c {kind}: displacement rel L2 {ed:.2e}, force rel L2 {ef:.2e}")
    assert ed <= 1e-5
    assert ef <= 1e-4
    _, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)
    if kind == "ep":
        ee = rel_l2(g.edp(0), o.blocks[0].edp)
        say(f"fp32c ep: plastic extension rel L2 {ee:.2e}, {int((o.blocks[0].edp != 0).sum())} yielded bonds")
        assert np.count_nonzero(o.blocks[0].edp) > 100 and ee <= 1e-4


def test_fp32c_contact_plates(pd, orc):
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    o, g = run_both(pd, orc, spec, 30, nuf=5, prec="fp32c")
    y, v, f = g.download()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    ed = rel_l2(y - x, o.y - x)
    say(f"fp32c plates: displacement rel L2 {ed:.2e}")
    assert ed <= 1e-5


def test_graph_latch_poll_equals_eager(pd, monkeypatch):
    """Plates that touch after a dozen steps: the graph run reads the pending
    deformed latches back every 16 steps and captures its graphs again once the
    blocks have latched; the result equals the eager run bit for bit."""
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=1.0, dt=1e-4, approach_per_step=0.05, exponent=2)
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("PERIDYN_GRAPH", mode)
        g = pd.Env.from_spec(spec, precision="fp64")
        g.run(64, 5)
        runs[mode] = g.download()
        g.close()
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    y = runs["1"][0]
    # the plates have deformed (not only moved rigidly)
    d0 = (y - x)[: spec["blocks"][0]["x"].shape[0]]
    assert np.ptp(d0[:, 2]) > 1e-6
    for u, w in zip(runs["0"], runs["1"]):
        assert np.array_equal(u, w)


def test_graph_replay_equals_eager(pd, orc, monkeypatch):
    """Steps replayed from the captured HIP graphs equal eagerly launched steps
    bit for bit, over several pd_run calls (each capturing and replaying its own
    graphs: plain steps, and steps followed by the contact rebuild, replayed at
    every rebuild after the first; calls of at least 32 steps capture them);
    and both equal the oracle in the reference-order form."""
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    for prec in ("exact", "fp64"):
        runs = {}
        for mode in ("0", "1"):
            monkeypatch.setenv("PERIDYN_GRAPH", mode)
            g = pd.Env.from_spec(spec, precision=prec)
            snaps = []
            for n in (37, 42, 33):
                g.run(n, 5)
                snaps.append(g.download())
            runs[mode] = snaps
        for a, b in zip(runs["0"], runs["1"]):
            for u, w in zip(a, b):
                assert np.array_equal(u, w)
        if prec == "exact":
            o = orc.OracleEnv(spec)
            o.run(37, 5)
            o.run(42, 5, start=True)
            o.run(33, 5, start=True)
            assert np.array_equal(runs["1"][-1][0], o.y)
            assert np.array_equal(runs["1"][-1][2], o.f)


@pytest.mark.parametrize("kind", ["bb", "osb"])
@pytest.mark.parametrize("prec", ["exact", "fp64", "fp32c"])
def test_indentation_fixbc_against_oracle(pd, orc, kind, prec):
    """FixBC (FixBC.jl:46-65: y = last, v = 0 on the block's three bottom
    layers) together with a ToFroBC indenter that reverses after 80% of the
    steps (ToFroBC.jl:51-111) and a linear contact, in the shape of
    examples/indentation/block_indentor.jl: 1,600 steps, through contact and
    the reversal.  Bars: reference-order form bit-identical to the oracle; FP64
    form within 1e-10 relative L2 in displacement and force; FP32C within
    1e-5 in displacement; the fixed particles never move in any form; identical
    bond masks."""
    spec = W.indentation(2000, kind=kind, steps=1600)
    o, g = run_both(pd, orc, spec, 1600, 10, prec)
    y, v, f = g.download()
    fix = spec["bcs"][0]["mask"]
    x = np.concatenate([b["x"] for b in spec["blocks"]])
    assert np.array_equal(y[fix], x[fix]) and not np.any(v[fix])
    ed, ef = rel_l2(y - x, o.y - x), rel_l2(f, o.f)
    say(f"indentation {kind} {prec}: displacement {ed:.2e}, force {ef:.2e}, "
        f"indenter travel {float((o.y[~fix][:, 2] - x[~fix][:, 2]).min()):.3f}")
    if prec == "exact":
        assert np.array_equal(y, o.y) and np.array_equal(v, o.v) and np.array_equal(f, o.f)
    elif prec == "fp64":
        assert ed <= 1e-10 and ef <= 1e-10
    else:
        assert ed <= 1e-5
    assert np.array_equal(g.damage(), o.damage())
    g.close()


def test_contact_seam_per_model_against_oracle(pd, orc):
    """The per-model contact seam with the caller's positions:
    update_contact_neighs!(y, type, RM) (contacts.jl:434-476) through
    pd_contact_update + pd_contact_neighs, and short_range_contact!(y, f, ...)
    (contacts.jl:189-247) through pd_contact_force, each for one model of an env
    with two: the plates' own model and a second RM12 model on the reversed pair
    with a wider search and max_neighs 5, so its lists are truncated to the
    largest ids.  Positions: the plates pushed into each other and jittered.
    Bars: RM.neighs identical to the oracle's; f bit-identical to the oracle's
    f after the same model's terms are added to the same random f."""
    spec = W.plates_ep(nx=8, ny=8, nz=4, gap=0.5, dt=1e-4)
    spec["contacts"].append({"pair": ([2], [1]), "exponent": 1, "stiffness": 1234.5, "distanceD": 1.0,
                             "distanceX": 2.5, "max_neighs": 5})
    g = pd.Env.from_spec(spec, precision="exact")
    o = orc.OracleEnv(spec)
    rng = np.random.default_rng(7)
    n0 = len(spec["blocks"][0]["x"])
    y = o.y.copy()
    y[n0:, 2] -= 1.3
    y += rng.uniform(-0.05, 0.05, size=y.shape)
    for cidx in (0, 1):
        o.y = y.copy()
        o.update_neighs()
        g.contact_update(cidx, y)
        ref = o.contacts[cidx]["neighs"]
        got = g.contact_neighs(cidx)
        assert got.shape == ref.shape and np.array_equal(got, ref), cidx
        assert int((ref > 0).sum()) > 50
        f0 = rng.normal(size=y.shape)
        o.f = f0.copy()
        o._contact_force(o.contacts[cidx])
        out = g.contact_force(f0, cidx, y)
        assert np.array_equal(out, o.f), cidx
        assert not np.array_equal(out, f0)
    assert (o.contacts[1]["neighs"] > 0).sum(axis=1).max() == 5  # truncated
    g.close()


@pytest.mark.parametrize("prec", ["exact", "fp64"])
def test_rigidly_held_block_latch_and_upload(pd, orc, prec):
    """A block every particle of which one ToFroBC holds moves rigidly and never
    latches (its latch tests stop once it has read back unlatched); positions
    set from the host are tested again: a stretched upload of that block latches
    it at the next step, as the oracle's latch (general_material.jl:240-256)
    does.  A second, free block is pulled by grips meanwhile.  Bars: exact form
    bit-identical to the oracle, FP64 within 1e-10, over three calls, the
    upload and two more calls."""
    K = W.bulk_from_E(W.E0, W.NU0)
    p = dict(bulk_modulus=K, critical_stretch=W.S00, density=W.RHO0)
    xa = W.lattice_box(7, 7, 7)
    xb = W.lattice_box(5, 5, 5, origin=(20.0, 0.0, 0.0))
    spec = {"dt": 1e-4, "blocks": [W.block("bb", xa, W.DELTA, type_id=1, **p),
                                   W.block("bb", xb, W.DELTA, type_id=2, **p)], "contacts": []}
    n = len(xa) + len(xb)
    grips = W.grips_z(xa, 2, 20.0)
    for g_ in grips:
        g_["mask"] = np.concatenate([g_["mask"], np.zeros(len(xb), dtype=bool)])
    held = np.zeros(n, dtype=bool)
    held[len(xa):] = True
    spec["bcs"] = grips + [{"kind": "tofro", "mask": held, "rate": np.array([3.0, -1.0, 2.0]), "freq": 25.0,
                            "applyafter": 0}]
    o = orc.OracleEnv(spec)
    g = pd.Env.from_spec(spec, precision=prec)
    for k in (7, 20, 13):
        o.run(k, 1, start=True)
        g.run(k, 1)
    assert not o.blocks[1].deformed
    y = o.y.copy()
    c = y[len(xa):].mean(axis=0)
    y[len(xa):] = c + (y[len(xa):] - c) * np.array([1.0, 1.0, 1.002])
    o.y = y.copy()
    g.upload(y=y)
    for k in (5, 9):
        o.run(k, 1, start=True)
        g.run(k, 1)
    assert o.blocks[1].deformed
    yy, vv, ff = g.download()
    if prec == "exact":
        assert np.array_equal(yy, o.y) and np.array_equal(vv, o.v) and np.array_equal(ff, o.f)
    else:
        x = np.concatenate([xa, xb])
        assert rel_l2(yy - x, o.y - x) <= 1e-10 and rel_l2(ff, o.f) <= 1e-10
    assert np.any(ff[len(xa):] != 0.0)
    g.close()
