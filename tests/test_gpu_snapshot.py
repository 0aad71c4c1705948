"""Snapshot I/O from the device (save_state!, io.jl:44-54) and resident-state
checkpoints.

Bars: save_state columns equal the oracle's state after its own update_acc!
(bit-identical in the reference-order form: the text round-trips exactly); a
run continued from a checkpoint (same env or a fresh one built from the same
description) is bit-identical to the uninterrupted run; a checkpoint of another
shape is refused."""
import numpy as np
import pytest

from peridyn_amd import snapshot as S
from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from peridyn_amd import core
    return core


def test_save_state_matches_oracle(core, orc, tmp_path):
    spec = W.tensile_bb(n=6, dt=1e-3, pull_per_step=0.03)
    o = orc.OracleEnv(spec)
    o.run(9, 1)
    o.update_acc()
    g = core.Env.from_spec(spec, precision="exact")
    g.run(9, 1)
    fn = S.save_state(str(tmp_path / "env_1_step_9.data"), g)
    d = S.read_ovito(fn)
    assert np.array_equal(d["id"], np.arange(1, o.N + 1))
    assert np.array_equal(d["position"], o.y.T)
    assert np.array_equal(d["velocity"], o.v.T)
    assert np.array_equal(d["acceleration"], o.f.T)
    assert np.array_equal(d["mass"], o.mass)
    assert np.array_equal(d["volume"], o.volume)
    assert np.array_equal(d["type"], o.type)
    assert np.array_equal(d["damage"], o.damage())
    b = S.read_ovito(S.save_state_ovito_bc(str(tmp_path / "x.data"), g))
    want = np.full(o.N, 99)
    for i, bc in enumerate(spec["bcs"], start=1):
        want[bc["mask"]] = i
    assert np.array_equal(b["type"], want)


@pytest.mark.parametrize("kind,prec", [("bb", "exact"), ("ep", "fp64"), ("osb", "fp32c"), ("contact", "fp64")])
def test_checkpoint_resume_is_bit_identical(core, kind, prec):
    if kind == "contact":
        spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
        nuf = 5
    else:
        spec = W.tensile_bb(n=7, dt=1e-4, kind=kind, pull_per_step=0.004)
        if kind == "ep":
            spec["blocks"][0]["yield_stress"] = 20.0
        nuf = 1
    a = core.Env.from_spec(spec, precision=prec)
    a.run(13, nuf)
    blob = a.checkpoint()
    a.run(17, nuf)
    ref = a.download()
    b = core.Env.from_spec(spec, precision=prec)
    b.restore(blob)
    assert b.time_step == 13
    b.run(17, nuf)
    for u, w in zip(ref, b.download()):
        assert np.array_equal(u, w)
    assert np.array_equal(a.damage(), b.damage())
    if kind == "ep":
        assert np.array_equal(a.edp(0), b.edp(0))
    a.restore(blob)          # back into the original env
    a.run(17, nuf)
    for u, w in zip(ref, a.download()):
        assert np.array_equal(u, w)


def test_checkpoint_of_other_shape_is_refused(core):
    a = core.Env.from_spec(W.tensile_bb(n=6), precision="fp64")
    b = core.Env.from_spec(W.tensile_bb(n=7), precision="fp64")
    with pytest.raises(Exception):
        b.restore(a.checkpoint())


def test_force_density_after_odd_run_keeps_state(core, orc):
    """force_density_T! seam after an odd number of steps (the resident buffers
    have swapped), then the run continues: equal to the oracle."""
    spec = W.tensile_bb(n=6, dt=1e-4, pull_per_step=0.003)
    o = orc.OracleEnv(spec)
    o.run(3, 1)
    fo = np.zeros_like(o.y)
    o.blocks[0].force_density(o.y, fo)
    o.run(5, 1)
    g = core.Env.from_spec(spec, precision="exact")
    g.run(3, 1)
    y, v, f = g.download()
    fg = np.zeros_like(y)
    g.force_density(0, y, fg)
    assert np.array_equal(fg, fo)
    g.run(5, 1)
    y2, v2, f2 = g.download()
    assert np.array_equal(y2, o.y) and np.array_equal(v2, o.v) and np.array_equal(f2, o.f)
