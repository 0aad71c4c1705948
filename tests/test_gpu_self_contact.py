"""Self contact (ContactModel11: update_contact_neighs! contacts.jl:539-591,
short_range_contact! :357-380) on the GPU against the oracle's literal
restatement.

Scenario: a tensile block whose grips reverse every 20 steps, with bonds
broken from the first steps; distanceD = 3.2 so that the contact partner the
reference's merge walk can find (the largest-index family member, the +z bond
at 3 dx) is within equi_dist once its bond is broken.

Bars: reference-order form bit-identical (y, v, f, bond states); FP64 form
1e-10 relative L2; the contact terms change the trajectory."""
import math

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


def scenario(with_contact=True, exponent=2):
    spec = W.tensile_bb(n=8, dt=1e-4, pull_per_step=0.01, s0=0.003)
    for bc in spec["bcs"]:
        bc["freq"] = 20.0
        bc["applyafter"] = 0
    K = 18.0 * W.bulk_from_E(70e3) / (math.pi * 3.015 ** 4)
    if with_contact:
        spec["contacts"] = [{"pair": ([1],), "exponent": exponent, "stiffness": K, "distanceD": 3.2,
                             "distanceX": 2.0, "max_neighs": 50}]
    return spec


@pytest.mark.parametrize("exponent", [1, 2])
def test_self_contact_exact_bit_identical(core, orc, exponent):
    spec = scenario(exponent=exponent)
    o = orc.OracleEnv(spec)
    o.run(60, 5)
    assert np.count_nonzero(o.contacts[0]["neighs"]) > 0
    g = core.Env.from_spec(spec, precision="exact")
    g.run(60, 5)
    y, v, f = g.download()
    assert np.array_equal(y, o.y) and np.array_equal(v, o.v) and np.array_equal(f, o.f)
    _, intact = g.family(0)
    assert np.array_equal(intact, o.blocks[0].intact)


def test_self_contact_fp64_and_effect(core, orc):
    spec = scenario()
    o = orc.OracleEnv(spec)
    o.run(60, 5)
    g = core.Env.from_spec(spec, precision="fp64")
    g.run(60, 5)
    y, v, f = g.download()
    x = spec["blocks"][0]["x"]
    assert rel_l2(y - x, o.y - x) <= 1e-10 and rel_l2(f, o.f) <= 1e-10
    h = core.Env.from_spec(scenario(with_contact=False), precision="fp64")
    h.run(60, 5)
    yh, _, _ = h.download()
    assert rel_l2(yh - x, y - x) > 1e-6  # the self-contact terms act
