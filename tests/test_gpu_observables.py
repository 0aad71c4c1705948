"""Device observables (printdata columns, environment.jl:308-320) against the
oracle: logged rows of a run (step 1 and every `every`-th step, solvers.jl:158)
and the columns of the current state after update_acc!.

Bars (stated here): time identical; momentum and force sums within 1e-12
(reference-order form) / 1e-10 (FP64 form) of the oracle relative to the sum of
magnitudes (sums of cancelling terms have no relative scale of their own); mean
damage within 1e-14 (same bond states, different summation order).  Logging does
not perturb the run (bit-identical state with and without it)."""
import numpy as np
import pytest

from peridyn_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from peridyn_amd import core
    return core


def oracle_rows(o, nsteps, nuf, every):
    """OracleEnv.run with log_every, plus the sums of magnitudes |p|, |f| of
    each logged step (the scale of its cancelling sums)."""
    o.update_neighs()
    o.apply_bc_at0()
    rows, sp, sf = [], [], []
    for i in range(1, nsteps + 1):
        o.step()
        if i % every == 0 or i == 1:
            rows.append(o.printdata())
            sp.append(np.abs(o.p).sum())
            sf.append(np.abs(o.f).sum())
        if i % nuf == 0:
            o.update_neighs()
    return np.array(rows), np.array(sp)[:, None], np.array(sf)[:, None]


def check_rows(g, o, scale_p, scale_f, tol):
    assert g.shape == o.shape
    assert np.array_equal(g[:, 0], o[:, 0])
    assert np.all(np.abs(g[:, 1:4] - o[:, 1:4]) <= tol * scale_p)
    assert np.all(np.abs(g[:, 4:7] - o[:, 4:7]) <= tol * scale_f)
    assert np.all(np.abs(g[:, 7] - o[:, 7]) <= 1e-14)


@pytest.mark.parametrize("prec,tol", [("exact", 1e-12), ("fp64", 1e-10)])
def test_logged_tensile_with_breaking(core, orc, prec, tol):
    spec = W.tensile_bb(n=8, dt=1e-3, pull_per_step=0.02)
    o = orc.OracleEnv(spec)
    orows, scale_p, scale_f = oracle_rows(o, 40, 1, 10)
    g = core.Env.from_spec(spec, precision=prec)
    grows = g.run_logged(40, 1, 10)
    assert orows[-1][7] > 0  # bonds broke
    check_rows(grows, orows, scale_p, scale_f, tol)
    # logging does not perturb the state
    h = core.Env.from_spec(spec, precision=prec)
    h.run(40, 1)
    for a, b in zip(g.download(), h.download()):
        assert np.array_equal(a, b)


def test_logged_plates_contact_fp64(core, orc):
    spec = W.plates_ep(nx=8, ny=8, nz=5, gap=0.5, dt=1e-4, approach_per_step=0.05, exponent=2)
    o = orc.OracleEnv(spec)
    orows, scale_p, scale_f = oracle_rows(o, 30, 5, 7)
    g = core.Env.from_spec(spec, precision="fp64")
    grows = g.run_logged(30, 5, 7)
    assert grows.shape[0] == 5  # steps 1, 7, 14, 21, 28
    check_rows(grows, orows, scale_p, scale_f, 1e-10)


def test_observables_after_update_acc(core, orc):
    spec = W.tensile_bb(n=7, dt=1e-3, pull_per_step=0.01)
    o = orc.OracleEnv(spec)
    o.run(5, 1)
    o.update_acc()
    g = core.Env.from_spec(spec, precision="exact")
    g.run(5, 1)
    g.update_acc()
    got = g.observables()
    want = o.printdata()
    assert got[0] == want[0]
    assert np.all(np.abs(got[1:4] - want[1:4]) <= 1e-12 * np.abs(o.p).sum())
    assert np.all(np.abs(got[4:7] - want[4:7]) <= 1e-12 * np.abs(o.f).sum())
    assert abs(got[7] - want[7]) <= 1e-14


def test_batch_logged_rows_per_env(core, orc):
    specs = [W.tensile_bb(n=6, dt=1e-4), W.tensile_bb(n=7, dt=1e-4, kind="osb", pull_per_step=0.003)]
    envs = [core.Env.from_spec(s, precision="exact") for s in specs]
    log = core.run_batch_logged(envs, 20, 1, 5)
    assert log.shape == (5, 2, 8)
    for k, s in enumerate(specs):
        o = orc.OracleEnv(s)
        orows, scale_p, scale_f = oracle_rows(o, 20, 1, 5)
        check_rows(log[:, k, :], orows, scale_p, scale_f, 1e-12)
