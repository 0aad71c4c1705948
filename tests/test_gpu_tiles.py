"""Tiled passes (halos staged in LDS) against the untiled kernels of the same
form (FP64 and FP32-compensated) and against the oracle.

The tiled kernels evaluate every bond with the untiled kernels' expressions in
the same per-lane order, so for the same lanes per particle the results are
bit-identical (== on y, v, f, bond states, damage).  The internal particle
order differs (bricks vs. the input order), which the boundary translates: the
comparisons are in reference order.  Against the oracle the FP64 bar holds
(relative L2 <= 1e-10).
"""
import contextlib
import os
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


@contextlib.contextmanager
def tiles(on: bool):
    old = os.environ.get("PERIDYN_TILES")
    os.environ["PERIDYN_TILES"] = "1" if on else "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["PERIDYN_TILES"]
        else:
            os.environ["PERIDYN_TILES"] = old


def make(pd, spec, tiled, lanes, precision="fp64"):
    with tiles(tiled):
        env = pd.Env.from_spec(spec, precision=precision, lanes_per_particle=lanes)
    lay = env.layout()
    assert lay["tiled"] == tiled, lay
    assert lay["lanes_per_particle"] == lanes
    return env


def same_state(a, b):
    ya, va, fa = a.download()
    yb, vb, fb = b.download()
    assert np.array_equal(ya, yb)
    assert np.array_equal(va, vb)
    assert np.array_equal(fa, fb)
    assert np.array_equal(a.damage(), b.damage())
    for blk in range(len(a.spec["blocks"])):
        fa_, ia = a.family(blk)
        fb_, ib = b.family(blk)
        assert np.array_equal(fa_, fb_)
        assert np.array_equal(ia, ib)


def ep_tensile(n=9):
    spec = W.tensile_bb(n=n, dt=1e-4, kind="ep", pull_per_step=0.002)
    spec["blocks"][0]["yield_stress"] = 20.0
    return spec


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_tiled_equals_untiled(pd, kind, lanes):
    say("tiled vs untiled", kind, lanes)
    spec = ep_tensile() if kind == "ep" else W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.002)
    a = make(pd, spec, True, lanes)
    b = make(pd, spec, False, lanes)
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


def jittered(spec, scale=1e-3, seed=5):
    """Reference positions and volumes that are not FP32 values: the tiled
    passes stage full FP64 halo records instead of the compact ones."""
    rng = np.random.default_rng(seed)
    for b in spec["blocks"]:
        x = np.asarray(b["x"], dtype=np.float64)
        b["x"] = np.ascontiguousarray(x + rng.uniform(-scale, scale, size=x.shape) + 1e-9)
        b["vol"] = np.asarray(b["vol"], dtype=np.float64) * (1.0 + 1e-10)
    return spec


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
@pytest.mark.parametrize("lanes", [2, 4])
def test_tiled_full_records_equal_untiled(pd, kind, lanes):
    say("tiled full records vs untiled", kind, lanes)
    spec = ep_tensile() if kind == "ep" else W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.002)
    spec = jittered(spec)
    a = make(pd, spec, True, lanes)
    b = make(pd, spec, False, lanes)
    assert not a.layout()["compact_records"]
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


def test_compact_and_full_records_batched(pd):
    """One batch with a compact-record env and a full-record env: the set
    launches the full records for both; each result equals its run alone."""
    say("compact + full batch")
    s1 = ep_tensile()
    s2 = jittered(ep_tensile(n=8))
    alone = [make(pd, s1, True, 4), make(pd, s2, True, 4)]
    assert alone[0].layout()["compact_records"] and not alone[1].layout()["compact_records"]
    for e in alone:
        e.run(25, 1)
    mix = [make(pd, s1, True, 4), make(pd, s2, True, 4)]
    pd.run_batch(mix, 25, 1)
    for a, b in zip(alone, mix):
        same_state(a, b)
    for e in alone + mix:
        e.close()


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
def test_tiled_within_fp64_bar_of_oracle(pd, orc, kind):
    say("tiled vs oracle", kind)
    spec = ep_tensile() if kind == "ep" else W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.002)
    o = orc.OracleEnv(spec)
    o.run(40, 1)
    g = make(pd, spec, True, 4)
    g.run(40, 1)
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(f, o.f) <= 1e-10
    assert rel_l2(y - x, o.y - x) <= 1e-10
    g.close()


def test_tiled_bond_breaking_equals_untiled(pd):
    """Fast pull: bonds break during the run (the tiled kernels clear the packed
    intact bits together with the bond states)."""
    say("tiled breaking")
    spec = W.tensile_bb(n=10, dt=1e-3, kind="bb", pull_per_step=0.05)
    a = make(pd, spec, True, 4)
    b = make(pd, spec, False, 4)
    a.run(60, 1)
    b.run(60, 1)
    assert a.damage().max() > 0.0
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("lanes", [2, 4])
def test_tiled_contact_plates_equal_untiled(pd, lanes):
    say("tiled contact plates", lanes)
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    a = make(pd, spec, True, lanes)
    b = make(pd, spec, False, lanes)
    a.run(30, 5)
    b.run(30, 5)
    same_state(a, b)
    assert np.abs(a.download()[2]).max() > 0
    a.close()
    b.close()


def test_tiled_seams_then_run_equal_untiled(pd):
    """update_acc! through the untiled passes on a tiled env (bonds break
    there), then tiled steps: the packed intact bits follow the bond states."""
    say("tiled seams")
    spec = W.tensile_bb(n=10, dt=1e-3, kind="bb", pull_per_step=0.05)
    rng = np.random.default_rng(3)
    y0 = spec["blocks"][0]["x"] * 1.02 + rng.normal(scale=2e-2, size=spec["blocks"][0]["x"].shape)
    envs = []
    for tiled in (True, False):
        e = make(pd, spec, tiled, 4)
        e.update_acc(np.ascontiguousarray(y0))
        e.run(20, 1)
        envs.append(e)
    assert envs[0].damage().max() > 0.0
    same_state(envs[0], envs[1])
    for e in envs:
        e.close()


def test_tiled_checkpoint_resume_bit_identical(pd):
    say("tiled checkpoint")
    spec = ep_tensile()
    a = make(pd, spec, True, 4)
    a.run(30, 1)
    blob = a.checkpoint()
    a.run(20, 1)
    b = make(pd, spec, True, 4)
    b.restore(blob)
    b.run(20, 1)
    same_state(a, b)
    # a snapshot of the untiled layout does not load into a tiled env
    c = make(pd, spec, False, 4)
    from peridyn_amd._lib import PeriDynError
    with pytest.raises(PeriDynError):
        b.restore(c.checkpoint())
    for e in (a, b, c):
        e.close()


def test_tiled_and_untiled_batched_equal_alone(pd):
    say("tiled batch")
    s1 = ep_tensile()
    s2 = W.tensile_bb(n=8, dt=1e-4, kind="osb", pull_per_step=0.002)
    alone = [make(pd, s1, True, 4), make(pd, s2, False, 2)]
    for e in alone:
        e.run(25, 1)
    mix = [make(pd, s1, True, 4), make(pd, s2, False, 2)]
    pd.run_batch(mix, 25, 1)
    for a, b in zip(alone, mix):
        same_state(a, b)
    for e in alone + mix:
        e.close()


def test_fp32c_envs_with_ep_blocks_run_untiled(pd):
    """The FP32C elastoplastic passes are untiled kernels: an FP32C env holding
    an EP block is laid out untiled even where its halos would fit, while the
    same env in FP64 runs tiled."""
    spec = ep_tensile()
    e32 = pd.Env.from_spec(spec, precision="fp32c", lanes_per_particle=4)
    e64 = pd.Env.from_spec(spec, precision="fp64", lanes_per_particle=4)
    assert not e32.layout()["tiled"] and e64.layout()["tiled"]
    e32.close()
    e64.close()


@pytest.mark.parametrize("kind", ["bb", "osb"])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_tiled_fp32c_equals_untiled(pd, kind, lanes):
    """FP32-compensated form: BB / OSB blocks run the tiled FP32C kernels,
    bit-identical to their untiled kernels."""
    say("tiled fp32c vs untiled", kind, lanes)
    spec = W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.002)
    a = make(pd, spec, True, lanes, "fp32c")
    b = make(pd, spec, False, lanes, "fp32c")
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


def test_tiled_fp32c_breaking_and_batch(pd):
    say("tiled fp32c breaking + batch")
    s1 = W.tensile_bb(n=10, dt=1e-3, kind="bb", pull_per_step=0.05)
    s2 = W.tensile_bb(n=8, dt=1e-4, kind="osb", pull_per_step=0.002)
    alone = [make(pd, s1, True, 4, "fp32c"), make(pd, s2, True, 2, "fp32c")]
    ref = [make(pd, s1, False, 4, "fp32c"), make(pd, s2, False, 2, "fp32c")]
    for e in alone + ref:
        e.run(50, 1)
    assert alone[0].damage().max() > 0.0
    for a, b in zip(alone, ref):
        same_state(a, b)
    mix = [make(pd, s1, True, 4, "fp32c"), make(pd, s2, False, 2, "fp32c")]
    pd.run_batch(mix, 50, 1)
    for a, b in zip(alone, mix):
        same_state(a, b)
    for e in alone + ref + mix:
        e.close()


@pytest.mark.parametrize("kind", ["bb", "ep"])
def test_device_and_host_tile_builds_agree(pd, kind):
    """The device halo build (sorted unique neighbour ids per tile, packed
    local indices) and the host build give the same tiles: identical runs."""
    say("tile builds", kind)
    spec = ep_tensile() if kind == "ep" else W.tensile_bb(n=10, dt=1e-3, kind="bb", pull_per_step=0.05)
    a = make(pd, spec, True, 4)
    old = os.environ.get("PERIDYN_TILE_BUILD")
    os.environ["PERIDYN_TILE_BUILD"] = "host"
    try:
        b = make(pd, spec, True, 4)
    finally:
        if old is None:
            del os.environ["PERIDYN_TILE_BUILD"]
        else:
            os.environ["PERIDYN_TILE_BUILD"] = old
    assert a.layout() == b.layout()
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("kind,horizon,stride", [("bb", 2.6, "608"), ("bb", 3.8, "896"), ("bb", 4.6, "run time"),
                                                 ("osb", 2.6, "608"), ("osb", 3.8, "896")])
def test_halo_stride_classes_equal_untiled(pd, kind, horizon, stride):
    """The tiled passes use a compile-time LDS stride when the set's largest halo
    fits 608 or 896 records, else the halo capacity at run time; each class
    gives the untiled kernels' bits.  (An OSB block with halos beyond 896 runs
    untiled: its records exceed the LDS budget.)"""
    say("stride class", kind, horizon, stride)
    spec = W.tensile(3000, horizon=horizon, kind=kind, pull_per_step=0.002)
    a = make(pd, spec, True, 4)
    hmax = a.layout()["max_halo"]
    assert (hmax <= 608) if stride == "608" else (608 < hmax <= 896) if stride == "896" else (hmax > 896), hmax
    b = make(pd, spec, False, 4)
    a.run(30, 1)
    b.run(30, 1)
    same_state(a, b)
    for e in (a, b):
        e.close()


def moved(spec, shift=(0.0, 0.0, 0.0), scale=1.0):
    for b in spec["blocks"]:
        x = np.asarray(b["x"], dtype=np.float64) * scale + np.asarray(shift)
        b["x"] = np.ascontiguousarray(x)
        if scale != 1.0:
            b["vol"] = np.asarray(b["vol"], dtype=np.float64) * scale ** 3
            b["horizon"] = float(b["horizon"]) * scale
    return spec


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
@pytest.mark.parametrize("case", ["far", "negative", "half", "volumes"])
def test_tiled_lattice_positions_equal_untiled(pd, kind, case):
    """Compact records with and without the integer-lattice table of |X|: an
    integer lattice far from the origin (its coordinates wrap the table's 16-bit
    lattice words), one at negative coordinates, a half-integer lattice (FP32
    values, general |X|), and an integer lattice of unequal FP32 volumes (the
    volumes from the records, not the block's one volume).  Each equals the
    untiled kernels bit for bit."""
    say("lattice positions", kind, case)
    spec = ep_tensile() if kind == "ep" else W.tensile_bb(n=9, dt=1e-4, kind=kind, pull_per_step=0.002)
    if case == "far":
        spec = moved(spec, shift=(40000.0, -20011.0, 123457.0))
    elif case == "negative":
        spec = moved(spec, shift=(-37.0, -8.0, -1000.0))
    elif case == "volumes":
        for b in spec["blocks"]:
            v = np.asarray(b["vol"], dtype=np.float64)
            b["vol"] = np.ascontiguousarray(v * (1.0 + (np.arange(v.size) % 7) / 8.0))
    else:
        spec = moved(spec, scale=0.5)
    a = make(pd, spec, True, 4)
    assert a.layout()["compact_records"]
    b = make(pd, spec, False, 4)
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
@pytest.mark.parametrize("build", ["device", "host"])
def test_partial_brick_tiles_equal_untiled(pd, kind, build):
    """A 12 x 12 x 10 slab: its top brick layer holds 2 of a brick's 4 chunk
    layers, and x-neighbouring partial bricks share one tile.  Tiled (device or
    host tile build) equals untiled bit for bit."""
    say("partial bricks", kind, build)
    x = W.lattice_box(12, 12, 10)
    K = W.bulk_from_E(W.E0, W.NU0)
    params = dict(bulk_modulus=K, critical_stretch=W.S00, density=W.RHO0)
    if kind != "bb":
        params["shear_modulus"] = 3.0 * K * (1 - 2 * W.NU0) / (2 * (1 + W.NU0))
    if kind == "ep":
        params["yield_stress"] = 20.0
    spec = {"dt": 1e-4, "blocks": [W.block(kind, x, W.DELTA, **params)], "contacts": [],
            "bcs": W.grips_z(x, 2, 0.002 / 1e-4)}
    old = os.environ.get("PERIDYN_TILE_BUILD")
    if build == "host":
        os.environ["PERIDYN_TILE_BUILD"] = "host"
    try:
        a = make(pd, spec, True, 4)
    finally:
        if old is None:
            os.environ.pop("PERIDYN_TILE_BUILD", None)
        else:
            os.environ["PERIDYN_TILE_BUILD"] = old
    b = make(pd, spec, False, 4)
    a.run(40, 1)
    b.run(40, 1)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_forked_bond_based_passes_equal_one_stream(pd, precision):
    """A batch of bond-based, ordinary state-based and elastoplastic specimens:
    the bond-based force passes run on a second stream beside the dilatation
    passes (PERIDYN_FORK=1 forks whatever the sizes; unset, sets whose
    launches are short fork); the results equal the one-stream run bit for
    bit, graph-replayed and eager."""
    say("fork", precision)
    specs = [W.tensile_bb(n=9, dt=1e-4, kind="bb", pull_per_step=0.002),
             W.tensile_bb(n=10, dt=1e-4, kind="osb", pull_per_step=0.002), ep_tensile(8),
             W.tensile_bb(n=11, dt=1e-4, kind="bb", pull_per_step=0.003)]
    out = {}
    for mode in ("0", "1"):
        old = os.environ.get("PERIDYN_FORK")
        os.environ["PERIDYN_FORK"] = mode
        try:
            envs = [pd.Env.from_spec(s, precision=precision, lanes_per_particle=4) for s in specs]
            pd.run_batch(envs, 40, 1)   # graph-replayed steps
            pd.run_batch(envs, 7, 1)    # eager (short call)
            out[mode] = envs
        finally:
            if old is None:
                os.environ.pop("PERIDYN_FORK", None)
            else:
                os.environ["PERIDYN_FORK"] = old
    for a, b in zip(out["0"], out["1"]):
        same_state(a, b)
    for e in out["0"] + out["1"]:
        e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_cost_ordered_tiles_equal_brick_order(pd, precision):
    """Specimens of different horizons and sizes stepped together (tiles of
    unequal cost in one launch, partial bricks among them): with the tiles of
    every launch ordered by cost and dealt over the XCDs (PERIDYN_TILE_ORDER=1)
    the results equal the brick-ordered run bit for bit, graph-replayed and
    eager; one indentation request adds its contact pass."""
    say("tile order", precision)
    specs = [W.tensile(3000, horizon=3.9, kind="bb"), W.tensile(5000, horizon=2.6, kind="osb"),
             W.tensile(1200, horizon=3.1, kind="osb"), W.tensile(7000, horizon=2.8, kind="bb"),
             W.indentation(2000, horizon=3.0, kind="bb", steps=60)]
    out = {}
    for mode in ("0", "1"):
        old = os.environ.get("PERIDYN_TILE_ORDER")
        os.environ["PERIDYN_TILE_ORDER"] = mode
        try:
            envs = [pd.Env.from_spec(s, precision=precision, lanes_per_particle=4) for s in specs]
            pd.run_batch(envs, 40, 10)  # graph-replayed steps
            pd.run_batch(envs, 7, 10)   # eager (short call)
            out[mode] = envs
        finally:
            if old is None:
                os.environ.pop("PERIDYN_TILE_ORDER", None)
            else:
                os.environ["PERIDYN_TILE_ORDER"] = old
    for a, b in zip(out["0"], out["1"]):
        assert a.layout()["tiled"]
        same_state(a, b)
    for e in out["0"] + out["1"]:
        e.close()
