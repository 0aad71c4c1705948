"""QSDrag quasi-static relaxation (minimize!, src/solvers/minimize.jl:105-231)
on the GPU against the oracle.

Bars: reference-order form bit-identical (positions, forces, bond states,
iteration count and final maxima); FP64 form positions within 1e-10 relative
L2 of the oracle, same iteration count."""
import numpy as np
import pytest

from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from peridyn_amd import core
    return core


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def pulled_block(kind="bb"):
    spec = W.tensile_bb(n=8, dt=1e-4, pull_per_step=0.002, kind=kind)
    if kind == "ep":
        spec["blocks"][0]["yield_stress"] = 20.0
    return spec


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
def test_minimize_exact_bit_identical(core, orc, kind):
    spec = pulled_block(kind)
    o = orc.OracleEnv(spec)
    o.run(20, 1)
    ores = o.minimize(1e-5, 100.0, max_iter=60, x_tol=1e-12, f_tol=1e-6)
    g = core.Env.from_spec(spec, precision="exact")
    g.run(20, 1)
    gres = g.minimize(1e-5, 100.0, max_iter=60, x_tol=1e-12, f_tol=1e-6)
    assert np.isfinite(ores[1]) and np.isfinite(ores[2])
    assert gres == ores
    y, v, f = g.download()
    assert np.array_equal(y, o.y) and np.array_equal(f, o.f) and np.array_equal(v, o.v)
    assert g.time_step == o.time_step == 21
    fam, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)
    if kind == "ep":
        assert np.array_equal(g.edp(0), o.blocks[0].edp)


def test_minimize_fp64_iterations(core, orc):
    spec = pulled_block("bb")
    o = orc.OracleEnv(spec)
    o.run(20, 1)
    o.minimize(1e-5, 100.0, max_iter=300, x_tol=1e-12, f_tol=1e-6)
    g = core.Env.from_spec(spec, precision="fp64")
    g.run(20, 1)
    it, fm, xm = g.minimize(1e-5, 100.0, max_iter=300, x_tol=1e-12, f_tol=1e-6)
    assert it == 300
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(y - x, o.y - x) <= 1e-10
    # the dynamics go on from the relaxed state
    g.run(5, 1)
    o.run(5, 1)
    y, v, f = g.download()
    assert rel_l2(y - x, o.y - x) <= 1e-10


@pytest.mark.parametrize("prec", ["exact", "fp64"])
def test_minimize_converges_at_the_oracle_iteration(core, orc, prec):
    """A stretched bond pair relaxes to its rest length; the stopping iteration
    (first with max|f| < f_tol and max|dy| < x_tol) is the oracle's."""
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    b = W.block("bb", x, 3.015, bulk_modulus=W.bulk_from_E(70e3), critical_stretch=1.0, density=2.7e-3)
    b["y"] = x.copy()
    b["y"][1, 0] += 0.004
    spec = {"dt": 1e-3, "blocks": [b], "contacts": [], "bcs": []}
    o = orc.OracleEnv(spec)
    ores = o.minimize(1e-4, 300.0, max_iter=20000, x_tol=1e-9, f_tol=1e-3)
    g = core.Env.from_spec(spec, precision=prec)
    gres = g.minimize(1e-4, 300.0, max_iter=20000, x_tol=1e-9, f_tol=1e-3)
    assert gres[0] == ores[0] < 20000
    y, v, f = g.download()
    if prec == "exact":
        assert gres == ores and np.array_equal(y, o.y)
    else:
        assert np.abs(y - o.y).max() <= 1e-12


def test_minimize_onlyatstart_bc_against_oracle(core, orc):
    """An onlyatstart grip (minimize.jl:138-143): applied once, then relaxed
    with the free particles; reference-order form bit-identical to the oracle
    (positions, forces, iterations), and different from the held grip."""
    spec = pulled_block("bb")
    spec["bcs"][1]["onlyatstart"] = True
    o = orc.OracleEnv(spec)
    o.run(20, 1)
    ores = o.minimize(1e-5, 100.0, max_iter=40, x_tol=1e-12, f_tol=1e-6)
    g = core.Env.from_spec(spec, precision="exact")
    g.run(20, 1)
    gres = g.minimize(1e-5, 100.0, max_iter=40, x_tol=1e-12, f_tol=1e-6)
    assert gres == ores
    y, v, f = g.download()
    assert np.array_equal(y, o.y) and np.array_equal(f, o.f)
    held = core.Env.from_spec(pulled_block("bb"), precision="exact")
    held.run(20, 1)
    held.minimize(1e-5, 100.0, max_iter=40, x_tol=1e-12, f_tol=1e-6)
    top = spec["bcs"][1]["mask"]
    assert not np.array_equal(held.download()[0][top], y[top])


def test_bc_dims_subset_is_refused(core):
    """dims selecting all three dimensions is accepted; a subset is refused with
    PD_E_ARG, where the reference raises DimensionMismatch."""
    from peridyn_amd import _lib
    spec = pulled_block("bb")
    spec["bcs"][0]["dims"] = [True, True, True]
    core.Env.from_spec(spec, precision="fp64").close()
    spec["bcs"][0]["dims"] = [False, False, True]
    with pytest.raises(_lib.PeriDynError) as ei:
        core.Env.from_spec(spec, precision="fp64")
    assert ei.value.code == _lib.PD_E_ARG
