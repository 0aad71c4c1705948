"""Pins the CPU oracle (oracle/) against closed forms, invariants and brute force.

The reference ships no tests, fixtures or golden vectors and cannot run here
(Julia and its PDMaterialPoints path dependency are absent), so every oracle
function is pinned by mathematics instead, chosen so that a dropped term, a wrong
sign/index or a transposed operand fails at least one check:

  cal_family        brute-force O(n^2) set equality; 122-neighbour interior shell
                    count at delta = 3.015; 816,728 bonds on the 20^3 tensile block;
                    descending rows + zero pad row; BoundsError cases
  weighted_volume   interior closed form sum |xi| = 285.2214799927738; brute force
  dilatation        homogeneous stretch => theta = 3 eps; translation => 0
  BB force          two-particle closed form c s vol/rho Yhat; momentum balance;
                    interior zero force under homogeneous stretch; break threshold
  OSB force         two-particle closed form 18 K e / r^2 Yhat / rho
  EP force          large-yield limit equals OSB; radial return lands on the yield
                    surface (recomputed td_norm2 == 2 psi); td_norm2 = 0 => elastic
  Verlet / BCs      free flight, MoveBC accumulation, FixBC
  contact           two-particle closed form; search == brute force + truncation
  julia_pow_int     exact small powers; correctly rounded x^4, x^5
  deformed latch    translated blocks skip forces, perturbed blocks do not
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from peridyn_amd import workloads as W


def brute_family(x, h):
    n = x.shape[0]
    out = []
    for i in range(n):
        # reference predicate operand order: (a1-a2)*(a1-a2) + ... with 1 = i, 2 = j
        a = x[i][None, :] - x
        r2 = (a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]) + a[:, 2] * a[:, 2]
        s = set(np.nonzero(r2 < h * h)[0].tolist()) - {i}
        out.append(s)
    return out


def family_sets(fam):
    return [set((fam[i][fam[i] > 0] - 1).tolist()) for i in range(fam.shape[0])]


# ---------------------------------------------------------------------------
def test_family_matches_brute_force_random_cloud(orc):
    rng = np.random.default_rng(7)
    x = rng.uniform(0.0, 9.0, size=(500, 3))
    fam, intact = orc.cal_family(x, 1.7, 200)
    assert family_sets(fam) == brute_family(x, 1.7)
    # rows descending, zero padded, at least one zero row, intact == fam > 0
    for i in range(fam.shape[0]):
        row = fam[i]
        nz = row[row > 0]
        assert np.all(np.diff(nz) < 0)
        assert np.all(row[len(nz):] == 0)
    assert fam.shape[1] == int((fam > 0).sum(axis=1).max()) + 1
    assert np.array_equal(intact, (fam > 0).astype(np.uint8))


def test_family_lattice_shell_counts(orc):
    x = W.lattice_box(11, 11, 11)
    fam, _ = orc.cal_family(x, 3.015, 200)
    deg = (fam > 0).sum(axis=1)
    center = 5 + 11 * (5 + 11 * 5)
    assert deg[center] == 122  # shells |xi|^2 in {1:6,2:12,3:8,4:6,5:24,6:24,8:12,9:30}
    # count by shell
    offs = x[np.array(sorted(family_sets(fam)[center]))] - x[center]
    sq = (offs ** 2).sum(axis=1).astype(int)
    counts = {k: int((sq == k).sum()) for k in sorted(set(sq))}
    assert counts == {1: 6, 2: 12, 3: 8, 4: 6, 5: 24, 6: 24, 8: 12, 9: 30}
    assert family_sets(fam) == brute_family(x, 3.015)


def test_family_tensile_block_bond_count(orc):
    x = W.lattice_box(20, 20, 20)
    fam, intact = orc.cal_family(x, 3.015, 200)
    assert int(intact.sum()) == 816728
    assert fam.shape[1] == 123


def test_family_overflow_is_bounds_error(orc):
    x = W.lattice_box(9, 9, 9)
    with pytest.raises(IndexError):
        orc.cal_family(x, 3.015, 122)  # max degree 122 == M -> end:-1:0
    with pytest.raises(IndexError):
        orc.cal_family(x, 3.015, 100)  # degree > M while filling
    fam, _ = orc.cal_family(x, 3.015, 123)
    assert fam.shape[1] == 123


def test_weighted_volume_closed_form_and_brute_force(orc):
    x = W.lattice_box(11, 11, 11)
    fam, _ = orc.cal_family(x, 3.015, 200)
    m = orc.weighted_volume(x, np.ones(x.shape[0]), fam)
    center = 5 + 11 * (5 + 11 * 5)
    shells = {1: 6, 2: 12, 3: 8, 4: 6, 5: 24, 6: 24, 8: 12, 9: 30}
    closed = sum(c * math.sqrt(k) for k, c in shells.items())
    assert abs(closed - 285.2214799927738) < 1e-12
    assert abs(m[center] - closed) <= 1e-13 * closed
    sets = brute_family(x, 3.015)
    for i in [0, 17, 333, 1330]:
        ref = sum(float(np.linalg.norm(x[j] - x[i])) for j in sets[i])
        assert abs(m[i] - ref) <= 1e-12 * ref


def test_dilatation_homogeneous_and_rigid(orc):
    x = W.lattice_box(8, 8, 8)
    fam, intact = orc.cal_family(x, 3.015, 200)
    vol = np.ones(x.shape[0])
    m = orc.weighted_volume(x, vol, fam)
    eps = 1e-3
    th = orc.dilatation(fam, intact, x, (1.0 + eps) * x, vol, m)
    assert np.allclose(th, 3 * eps, rtol=1e-10, atol=0)
    th0 = orc.dilatation(fam, intact, x, x + np.array([0.5, -0.25, 2.0]), vol, m)
    assert np.all(th0 == 0.0)


# ---------------------------------------------------------------------------
def two_particle_spec(kind, r=1.5, stretch=1.002, rho=2.0, K=3.0, G=1.2, s0=0.5, sy=1e30):
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    y = np.array([[0.0, 0.0, 0.0], [r * stretch, 0.0, 0.0]])
    b = W.block(kind, x, 2.0, y=y, bulk_modulus=K, shear_modulus=G, critical_stretch=s0,
                density=rho, yield_stress=sy)
    return {"dt": 1.0, "blocks": [b], "contacts": [], "bcs": []}


def test_bb_two_particle_closed_form(orc):
    r, st, rho, K = 1.5, 1.002, 2.0, 3.0
    env = orc.OracleEnv(two_particle_spec("bb", r, st, rho, K))
    env.update_acc()
    c = 18 * K / math.pi / 2.0 ** 4
    s = (r * st - r) / r
    fx = c * s * 1.0 / rho
    assert np.allclose(env.f[0], [fx, 0, 0], rtol=1e-13, atol=0)
    assert np.allclose(env.f[1], [-fx, 0, 0], rtol=1e-13, atol=0)


def test_osb_two_particle_closed_form(orc):
    r, st, rho, K, G = 1.5, 1.002, 2.0, 3.0, 1.2
    env = orc.OracleEnv(two_particle_spec("osb", r, st, rho, K, G))
    env.update_acc()
    e = r * st - r
    fx = 18 * K * e / r ** 2 / rho
    assert np.allclose(env.f[0], [fx, 0, 0], rtol=1e-12, atol=0)
    assert np.allclose(env.f[1], [-fx, 0, 0], rtol=1e-12, atol=0)


def test_ep_large_yield_equals_osb(orc):
    spec_o = W.tensile_bb(n=7, kind="osb")
    spec_e = W.tensile_bb(n=7, kind="ep")
    spec_e["blocks"][0]["yield_stress"] = 1e30
    rng = np.random.default_rng(3)
    y = spec_o["blocks"][0]["x"] * 1.001 + rng.normal(scale=1e-3, size=spec_o["blocks"][0]["x"].shape)
    for s in (spec_o, spec_e):
        s["blocks"][0]["y"] = y.copy()
    eo, ee = orc.OracleEnv(spec_o), orc.OracleEnv(spec_e)
    eo.update_acc(); ee.update_acc()
    assert np.allclose(ee.f, eo.f, rtol=1e-9, atol=1e-9 * np.abs(eo.f).max())
    assert np.all(ee.blocks[0].edp == 0.0)


def test_ep_radial_return_lands_on_yield_surface(orc):
    # simple shear gives deviatoric bond extension; a low yield stress makes
    # particles yield.  For every yielded particle i all its bonds share
    # C = sqrt(2 psi / td_norm2_i), so the returned state satisfies
    # sum_j (15G (ed - edp_new) w_m)^2 vol_j == 2 psi.
    import ctypes
    def make(sy):
        spec = W.tensile_bb(n=5, kind="ep")
        b0 = spec["blocks"][0]
        b0["yield_stress"] = sy
        b0["y"] = b0["x"] + np.stack([0.01 * b0["x"][:, 2], 0 * b0["x"][:, 0], 0 * b0["x"][:, 0]], axis=1)
        spec["bcs"] = []
        return orc.OracleEnv(spec)

    probe = make(1e30)
    probe.update_acc()
    med = float(np.median(probe.blocks[0].td_norm2))
    sy = math.sqrt(med / 2 * 3.015 ** 5 * 8 * math.pi / 75)  # 2 psi(sy) == median td_norm2
    env = make(sy)
    b = env.blocks[0]
    env.update_acc()
    td_before = b.td_norm2.copy()
    psi = b.psi[0]
    yielded = td_before >= 2 * psi
    assert yielded.sum() > 10 and (~yielded).sum() > 0
    assert np.all(b.edp[~yielded] == 0.0) and np.any(b.edp[yielded] != 0.0)
    orc.lib().orc_td_norm2(ctypes.byref(b._s), orc._p(env.y))
    assert np.allclose(b.td_norm2[yielded], 2 * psi, rtol=1e-10, atol=0)
    assert np.array_equal(b.td_norm2[~yielded], td_before[~yielded])


def test_ep_zero_td_norm_is_elastic(orc):
    # undeformed-but-latched block: td_norm2 = 0 -> C = Inf -> elastic, edp stays 0
    spec = two_particle_spec("ep", stretch=1.0 + 1e-12, sy=1.0)
    env = orc.OracleEnv(spec)
    env.blocks[0].deformed = True
    env.y[1, 0] = env.blocks[0].x[1, 0]
    env.update_acc()
    assert np.all(env.blocks[0].edp == 0.0)
    assert np.all(env.f == 0.0)


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
def test_momentum_balance_and_homogeneous_interior(orc, kind):
    spec = W.tensile_bb(n=9, kind=kind)
    b = spec["blocks"][0]
    rng = np.random.default_rng(11)
    b["y"] = b["x"] + rng.normal(scale=2e-3, size=b["x"].shape)
    env = orc.OracleEnv(spec)
    env.update_acc()
    P = (env.mass[:, None] * env.f).sum(axis=0)
    assert np.all(np.abs(P) <= 1e-10 * (env.mass[:, None] * np.abs(env.f)).sum(axis=0))
    # homogeneous stretch: particles whose force stencil sees only full
    # neighbourhoods feel ~0 (BB: one horizon from every face; OSB/EP: two,
    # because t_ij uses m_j and theta_j of the neighbours)
    n2, lo = (9, 3) if kind == "bb" else (16, 7)
    spec2 = W.tensile_bb(n=n2, kind=kind)
    spec2["blocks"][0]["y"] = spec2["blocks"][0]["x"] * 1.001
    env2 = orc.OracleEnv(spec2)
    env2.update_acc()
    x = spec2["blocks"][0]["x"]
    interior = np.all((x >= lo) & (x <= n2 - 1 - lo), axis=1)
    assert interior.sum() >= 8
    scale = np.abs(env2.f).max()
    assert scale > 0
    assert np.abs(env2.f[interior]).max() <= 1e-9 * scale


def test_bb_break_threshold(orc):
    s0 = 0.01
    for st, broken in [(1.0 + s0 * 0.999, False), (1.0 + s0 * 1.001, True), (0.5, False)]:
        spec = two_particle_spec("bb", r=1.0, stretch=st, s0=s0)
        env = orc.OracleEnv(spec)
        env.update_acc()
        assert bool(env.blocks[0].intact[0, 0] == 0) == broken
        if broken:
            assert np.all(env.f == 0.0)


# ---------------------------------------------------------------------------
def test_verlet_free_flight_and_bcs(orc):
    x = W.lattice_box(1, 1, 3, dx=10.0)  # far apart: no bonds with horizon 3
    b = W.block("bb", x, 3.0, v=(0.0, 0.0, 0.0), bulk_modulus=1.0, critical_stretch=1.0, density=1.0)
    b["v"] = np.array([[1.0, 2.0, -1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    fix = np.array([False, True, False])
    mv = np.array([False, False, True])
    spec = {"dt": 0.5, "blocks": [b], "contacts": [],
            "bcs": [{"kind": "fix", "mask": fix},
                    {"kind": "tofro", "mask": mv, "rate": np.array([0.0, 0.25, 1.0]), "freq": math.inf, "applyafter": 0}]}
    env = orc.OracleEnv(spec)
    env.run(8)
    assert np.array_equal(env.y[0], x[0] + 8 * 0.5 * np.array([1.0, 2.0, -1.0]))
    assert np.array_equal(env.y[1], x[1]) and np.all(env.v[1] == 0.0)
    assert np.array_equal(env.y[2], x[2] + 8 * (0.5 * np.array([0.0, 0.25, 1.0])))
    assert np.array_equal(env.v[2], [0.0, 0.25, 1.0])


def test_tofro_direction_flip(orc):
    x = W.lattice_box(1, 1, 1)
    b = W.block("bb", x, 3.0, bulk_modulus=1.0, critical_stretch=1.0, density=1.0)
    spec = {"dt": 1.0, "blocks": [b], "contacts": [],
            "bcs": [{"kind": "tofro", "mask": np.array([True]), "rate": np.array([1.0, 0.0, 0.0]),
                     "freq": 3, "applyafter": 0}]}
    env = orc.OracleEnv(spec)
    zs = []
    for _ in range(9):
        env.run(1, start=(env.time_step == 0))
        zs.append(env.y[0, 0])
    # check! flips after steps whose pre-increment time_step is 3, 6 (t > 0 and t % 3 == 0)
    assert zs == [1, 2, 3, 4, 3, 2, 1, 2, 3]


# ---------------------------------------------------------------------------
def test_contact_two_particle_closed_form(orc):
    for expo in (1, 2):
        xa = np.array([[0.0, 0.0, 0.0]])
        xb = np.array([[0.0, 0.0, 0.7]])
        ba = W.block("bb", xa, 3.0, type_id=1, bulk_modulus=1.0, critical_stretch=1.0, density=2.0)
        bb = W.block("bb", xb, 3.0, type_id=2, bulk_modulus=1.0, critical_stretch=1.0, density=4.0)
        K = 10.0
        spec = {"dt": 1.0, "blocks": [ba, bb], "bcs": [],
                "contacts": [{"pair": ([1], [2]), "exponent": expo, "stiffness": K,
                              "distanceD": 0.9, "distanceX": 2.0, "max_neighs": 5}]}
        env = orc.OracleEnv(spec)
        env.update_neighs()
        env.update_acc()
        e, r = 0.9, 0.7
        s = K * ((e - r) / e) ** expo / (r + 1e-10)
        F = s * np.array([0.0, 0.0, -0.7])
        assert np.allclose(env.f[0], F * (1.0 / 2.0), rtol=1e-14, atol=0)
        assert np.allclose(env.f[1], -F * (1.0 / 4.0), rtol=1e-14, atol=0)


def test_contact_exponent_value_selects_the_power(orc):
    """The contact models store the exponent as Float64 and Julia 1.10's
    Float64^Float64 takes power by squaring for integer-valued exponents: 3 and
    3.0 give the same bits (x*x*x), 2.5 takes the general power.  The binding's
    rule agrees with the oracle's."""
    from peridyn_amd import core
    for e in (1, 2, 3, 2.0, 3.0, -2.0, 4096.0):
        assert core.pow_by_squaring(e) and orc._julia_int_pow(e)
    for e in (1.5, 2.5, 4097.0, float("inf"), float("nan")):
        assert not core.pow_by_squaring(e) and not orc._julia_int_pow(e)
    forces = {}
    for expo in (3, 3.0, 2.5):
        xa = np.array([[0.0, 0.0, 0.0]])
        xb = np.array([[0.0, 0.0, 0.7]])
        ba = W.block("bb", xa, 3.0, type_id=1, bulk_modulus=1.0, critical_stretch=1.0, density=2.0)
        bb = W.block("bb", xb, 3.0, type_id=2, bulk_modulus=1.0, critical_stretch=1.0, density=4.0)
        spec = {"dt": 1.0, "blocks": [ba, bb], "bcs": [],
                "contacts": [{"pair": ([1], [2]), "exponent": expo, "stiffness": 10.0,
                              "distanceD": 0.9, "distanceX": 2.0, "max_neighs": 5}]}
        env = orc.OracleEnv(spec)
        env.update_neighs()
        env.update_acc()
        forces[expo] = env.f.copy()
    assert np.array_equal(forces[3], forces[3.0])
    strain = (0.9 - 0.7) / 0.9
    s3 = 10.0 * (strain * strain * strain) / (0.7 + 1e-10)
    assert forces[3][0, 2] == (s3 * -0.7) * (1.0 / 2.0)
    s25 = 10.0 * strain ** 2.5 / (0.7 + 1e-10)
    assert np.isclose(forces[2.5][0, 2], (s25 * -0.7) * (1.0 / 2.0), rtol=1e-14, atol=0)


def test_contact_general_power_within_one_ulp_of_exact(orc):
    """The general Float64 power of non-integer exponents (nonlinearspring.jl:140):
    Julia's `^(::Float64, ::Float64)` returns a value within one ulp of the exact
    x^y; the oracle's power is checked against x^y evaluated in 60-digit decimal
    arithmetic (exp(y ln x), rounded once to Float64) on strains across (0, 1]
    and exponents 0.5 .. 4.75: every value is within one ulp (and most are the
    correctly rounded one), so a wrong base, exponent or formula fails."""
    import decimal
    ctx = decimal.Context(prec=60)
    rng = np.random.default_rng(17)
    strains = np.concatenate([rng.uniform(1e-6, 1.0, 400), [1.0, 0.5, 0.1, 1e-3, 0.999999]])
    exact_hits = total = 0
    for e in (0.5, 0.75, 1.5, 2.25, 2.5, 3.3, 4.75):
        assert not orc._julia_int_pow(e)
        for x in strains:
            got = orc.contact_power(float(x), e)
            ex = ctx.exp(ctx.multiply(decimal.Decimal(e), ctx.ln(decimal.Decimal(float(x)))))
            want = float(ex)  # correctly rounded to Float64
            assert abs(got - want) <= np.spacing(want), (x, e, got, want)
            exact_hits += got == want
            total += 1
    assert exact_hits >= 0.99 * total
    # the integer-valued exponents take Julia's power by squaring instead
    assert orc.contact_power(0.3, 2.0) == orc.julia_pow_int(0.3, 2)


def test_contact_search_matches_brute_force(orc):
    rng = np.random.default_rng(5)
    x1 = rng.uniform(0, 6, size=(300, 3))
    x2 = rng.uniform(4, 10, size=(280, 3))
    d, M = 1.3, 6
    neighs = np.zeros((300, M), dtype=np.int64)
    orc.lib().orc_contact_neighs(300, orc._p(x1), 280, orc._p(x2), d, M, orc._p(neighs))
    for i in range(300):
        dd = ((x1[i] - x2) ** 2).sum(axis=1)
        cand = sorted((np.nonzero(dd < d * d)[0] + 1).tolist(), reverse=True)[:M]
        got = neighs[i][neighs[i] > 0].tolist()
        assert got == cand


def test_julia_pow_int(orc):
    rng = np.random.default_rng(1)
    for x in list(rng.uniform(0.1, 10.0, 200)) + [3.015, 0.5, 2.0]:
        x = float(x)
        assert orc.julia_pow_int(x, 1) == x
        assert orc.julia_pow_int(x, 2) == x * x
        assert orc.julia_pow_int(x, 3) == x * x * x
        for n in (4, 5):
            exact = float(Fraction(x) ** n)  # correctly rounded
            got = orc.julia_pow_int(x, n)
            assert abs(got - exact) <= math.ulp(exact)


def test_deformed_latch(orc):
    spec = W.tensile_bb(n=6)
    spec["bcs"] = []
    b = spec["blocks"][0]
    b["y"] = b["x"] + np.array([0.25, -3.0, 7.5])
    env = orc.OracleEnv(spec)
    env.blocks[0].deformed = False
    env.update_acc()
    assert not env.blocks[0].deformed and np.all(env.f == 0.0)
    # isapprox tolerance: rtol = sqrt(eps) relative to max(norm(a), norm(x))
    env.y[5, 2] += 1e-9
    env.update_acc()
    assert not env.blocks[0].deformed and np.all(env.f == 0.0)
    env.y[5, 2] += 1e-4
    env.update_acc()
    assert env.blocks[0].deformed and np.abs(env.f).max() > 0


def test_oracle_run_is_deterministic(orc):
    spec = W.tensile_bb(n=6, dt=1e-4)
    a, b = orc.OracleEnv(spec), orc.OracleEnv(spec)
    a.run(20); b.run(20)
    assert np.array_equal(a.y, b.y) and np.array_equal(a.f, b.f)


def test_printdata_closed_form_and_log_schedule(orc):
    """printdata columns (environment.jl:308-320) for a force-free block in
    uniform motion: time = n dt, p = (sum of vol*rho) v0, f = 0, damage 0; rows
    at step 1 and every 10th step (solvers.jl:158)."""
    spec = W.tensile_bb(n=5, dt=1e-3)
    spec["bcs"] = []
    b = spec["blocks"][0]
    v0 = np.array([0.5, -0.25, 2.0])
    b["v"] = np.broadcast_to(v0, b["x"].shape).copy()
    o = orc.OracleEnv(spec)
    rows = o.run(25, 1, log_every=10)
    assert rows.shape == (3, 8)
    M = float(np.sum(b["vol"] * b["density"]))
    for r, n in zip(rows, (1, 10, 20)):
        assert r[0] == n * 1e-3
        assert np.allclose(r[1:4], M * v0, rtol=1e-12, atol=0)
        assert np.all(r[4:7] == 0.0) and r[7] == 0.0


def test_printdata_damage_is_mean_of_cal_damage(orc):
    spec = W.tensile_bb(n=7, dt=1e-3, pull_per_step=0.05)
    o = orc.OracleEnv(spec)
    rows = o.run(12, 1, log_every=4)
    d = o.damage()
    assert d.max() > 0
    assert rows[-1][7] == d.sum() / d.size


def _pair_spec(stretch):
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    b = W.block("bb", x, 3.015, bulk_modulus=W.bulk_from_E(70e3), critical_stretch=1.0, density=2.7e-3)
    b["y"] = x.copy()
    b["y"][1, 0] += stretch
    return {"dt": 1e-3, "blocks": [b], "contacts": [], "bcs": []}


def test_qsdrag_first_iteration_closed_form(orc):
    """First minimize! iteration from opt = 0: force = f, dy = clip(0.5 f dt^2,
    ps/10), opt = dt f, left unclipped even where |opt| > ps/10/dt."""
    spec = _pair_spec(0.2)
    o = orc.OracleEnv(spec)
    o.update_acc()
    f0 = o.f.copy()
    y0 = o.y.copy()
    o2 = orc.OracleEnv(spec)
    dt = 0.5
    o2.minimize(dt, 1.0, max_iter=1)
    dy = np.clip(0.5 * f0 * (dt * dt), -0.1, 0.1)
    assert np.array_equal(o2.y, y0 + dy)
    assert np.abs(0.5 * f0 * dt * dt).max() > 0.1           # the clip was active
    assert np.array_equal(o2.qs_opt, dt * f0)                # opt not clipped
    assert np.abs(o2.qs_opt).max() > 0.1 / dt


def test_qsdrag_relaxes_pair_to_rest_length(orc):
    """Drag relaxation of a stretched bond pair ends at the rest length: with
    max|f| < f_tol the stretch left is below f_tol / (bond stiffness) ~ 1e-9."""
    o = orc.OracleEnv(_pair_spec(0.004))
    it, fm, xm = o.minimize(1e-4, 300.0, max_iter=20000, x_tol=1e-9, f_tol=1e-3)
    assert fm < 1e-3 and xm < 1e-9 and it < 20000
    sep = o.y[1] - o.y[0]
    assert abs(np.linalg.norm(sep) - 1.0) < 1e-8
    assert o.time_step == 1


def test_qsdrag_bc_particles_move_once(orc):
    spec = W.tensile_bb(n=8, dt=1e-3, pull_per_step=0.01)
    o = orc.OracleEnv(spec)
    o.apply_bc_at0()
    y0 = o.y.copy()
    o.minimize(1e-3, 10.0, max_iter=5)
    top = spec["bcs"][1]["mask"]
    assert np.allclose(o.y[top] - y0[top], [0.0, 0.0, 1e-3 * 10.0], rtol=0, atol=1e-15)


def test_qsdrag_onlyatstart_bc_particles_are_free(orc):
    """minimize.jl:138-143 masks only the particles of BCs without onlyatstart:
    an onlyatstart grip is applied once (moved by dt * rate) and then relaxes
    with the free particles.  One iteration from rest (opt = 0) moves a free
    particle by clip(0.5 f dt^2, ps/10), f the force at the BC-moved positions."""
    dt = 1e-3
    base = W.tensile_bb(n=8, dt=dt, pull_per_step=0.01)
    top = base["bcs"][1]["mask"]
    for only in (False, True):
        spec = W.tensile_bb(n=8, dt=dt, pull_per_step=0.01)
        spec["bcs"][1]["onlyatstart"] = only
        o = orc.OracleEnv(spec)
        o.apply_bc_at0()
        y0 = o.y.copy()
        moved = orc.OracleEnv(spec)
        moved.apply_bc_at0()
        for k, sgn in ((0, -1.0), (1, 1.0)):  # both grips applied once
            m = spec["bcs"][k]["mask"]
            moved.y[m] = y0[m] + np.array([0.0, 0.0, sgn * dt * 10.0])
        moved.update_acc()
        o.minimize(dt, 10.0, max_iter=1, x_tol=0.0, f_tol=0.0)
        want = y0[top] + np.array([0.0, 0.0, dt * 10.0])
        if only:
            a = o.blocks[0].particle_size / 10.0
            want = want + np.clip(0.5 * moved.f[top] * (dt * dt), -a, a)
            assert np.abs(moved.f[top]).max() > 0
        assert np.array_equal(o.y[top], want), only


def test_bc_dims_subset_is_a_dimension_mismatch(orc):
    """A BC on a subset of dimensions cannot take the reference's 3 x n update
    into y[dims, bool] (boundary_conditions.jl:108-110): the oracle refuses it
    as the reference raises DimensionMismatch; all three dims are accepted."""
    spec = W.tensile_bb(n=6)
    spec["bcs"][0]["dims"] = [True, True, True]
    orc.OracleEnv(spec)
    spec["bcs"][0]["dims"] = [False, False, True]
    with pytest.raises(ValueError):
        orc.OracleEnv(spec)


def _derived_rm11(y, fam, intact, equi):
    """The one-contact rule implied by the merge walk over descending columns:
    the largest family member, if that bond is broken and the pair is closer
    than equi (a different algorithm from the walk it is checked against)."""
    n = y.shape[0]
    out = np.zeros(n, dtype=np.int64)
    for i in range(n):
        j = fam[i, 0]
        if j > 0 and not intact[i, 0]:
            d = y[i] - y[j - 1]
            if ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) < equi * equi:
                out[i] = j
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rm11_merge_walk_keeps_only_the_largest_family_member(orc, seed):
    rng = np.random.default_rng(seed)
    x = W.lattice_box(6, 6, 6)
    fam, intact = orc.cal_family(x, 1.1, 40)  # faces: the largest member is +z at 1 dx
    intact = (intact.astype(bool) & (rng.uniform(size=intact.shape) < 0.5)).astype(np.uint8)
    y = x + rng.normal(scale=0.25, size=x.shape)
    nb = orc.rm11_neighs(y, fam, intact, 5.0, 1.0, 50)
    assert np.all(nb[:, 1:] == 0)                      # at most one contact each
    want = _derived_rm11(y, fam, intact, 1.0)
    assert np.array_equal(nb[:, 0], want)
    assert np.count_nonzero(want) > 5                  # the case occurs


def test_rm11_force_two_particle_closed_form(orc):
    y = np.array([[0.0, 0.0, 0.0], [0.6, 0.0, 0.0]])
    f = np.zeros_like(y)
    vol = np.array([1.0, 2.0])
    den = np.array([2.7e-3, 3.1e-3])
    nb = np.array([[2], [0]])
    K, equi = 100.0, 1.0
    orc.rm11_force(y, f, vol, den, nb, equi, K, 2.0, 1)
    strain = (equi - 0.6) / equi
    F = (K * strain * strain) / (0.6 + 1e-10) * (y[0] - y[1])
    assert np.allclose(f[0], F * (vol[1] / den[0]), rtol=1e-15, atol=0)
    assert np.allclose(f[1], -F * (vol[1] / den[1]), rtol=1e-15, atol=0)  # vol_j on both sides


@pytest.mark.parametrize("kind", ["bb", "osb", "ep"])
def test_threaded_timing_loops_match_single_thread(orc, kind):
    """The CPU baseline times the oracle with the reference's threaded loops on
    OpenMP threads.  Its pre-passes and bond states must equal the
    single-threaded oracle exactly, and its forces must equal them up to the
    rounding of the per-thread scatter sums, so the timed work is the same."""
    spec = W.tensile_bb(n=10, kind=kind)
    b = spec["blocks"][0]
    rng = np.random.default_rng(5)
    b["y"] = b["x"] * 1.002 + rng.normal(scale=4e-3, size=b["x"].shape)
    if kind == "ep":
        b["yield_stress"] = 1e-3  # yields: the EP update writes edp
    a, t = orc.OracleEnv(spec), orc.OracleEnv(spec)
    t.threaded = True
    for _ in range(3):  # edp carries over between passes
        a.update_acc()
        t.update_acc()
        ba, bt = a.blocks[0], t.blocks[0]
        assert np.array_equal(ba.intact, bt.intact)
        assert np.array_equal(ba.theta, bt.theta) and np.array_equal(ba.td_norm2, bt.td_norm2)
        assert np.array_equal(ba.edp, bt.edp)
        assert np.abs(a.f - t.f).max() <= 1e-12 * np.abs(a.f).max()
    assert (ba.intact == 0).sum() > (ba.fam == 0).sum()  # some bonds broke
    if kind == "ep":
        assert np.abs(ba.edp).max() > 0


def test_threaded_contact_search_matches_single_thread(orc):
    spec = W.plates_ep(nx=10, ny=10, nz=4, gap=0.5)
    a, t = orc.OracleEnv(spec), orc.OracleEnv(spec)
    t.threaded = True
    for e in (a, t):
        e.y[e.limits[1][0]:] -= [0.0, 0.0, 1.2]  # interpenetrate: many candidates
        e.update_neighs()
    assert all(np.array_equal(ca["neighs"], ct["neighs"]) for ca, ct in zip(a.contacts, t.contacts))
    assert sum(int((c["neighs"] > 0).sum()) for c in a.contacts) > 100


def test_cpu_baseline_leg_runs_threaded(orc):
    """bench.py's cpu_baseline leg on a small specimen: it steps the threaded
    oracle and reports the threads it used."""
    import bench
    spec = W.tensile_bb(n=8, kind="osb")
    env = orc.OracleEnv(spec)
    nb = int(sum((b.intact > 0).sum() for b in env.blocks))
    cb = bench.cpu_baseline(spec, 10, 0.2, nb)
    assert cb["unit"] == "bond-updates/s" and cb["kind"] == "port"
    assert cb["value"] > 0 and cb["cores"] >= 1
    assert f"{cb['cores']} OpenMP threads" in cb["sample"]


def test_cpu_baseline_batched_leg(orc):
    """bench.py's batched cpu_baseline leg: the threaded oracle on a sample of
    small specimens, with the sample stated."""
    import bench
    specs = [W.tensile(600, kind=k) for k in ("bb", "osb", "bb")]
    cb = bench.cpu_baseline_batched(specs, 3, 0.0)
    assert cb["unit"] == "bond-updates/s" and cb["kind"] == "port"
    assert cb["value"] > 0 and cb["cores"] >= 1
    assert cb["sample"].startswith("1 of the 3 specimens")


def test_bench_rejects_impossible_roofline():
    """A bench line whose dominant kernel claims more than the HBM peak, or
    more time per step than the step took, is rejected instead of printed."""
    import bench
    ok = {"roofline": {"frac": 0.4, "avg_launch_ms": 0.5, "launches": 20},
          "eager_profiled_run": {"ms_per_step": 1.0}}
    bench.check_line(ok, 20)
    for bad in ({"roofline": {"frac": 2.7, "avg_launch_ms": 0.01, "launches": 20},
                 "eager_profiled_run": {"ms_per_step": 0.1}},
                {"roofline": {"frac": 0.4, "avg_launch_ms": 0.5, "launches": 40},
                 "eager_profiled_run": {"ms_per_step": 0.6}}):
        with pytest.raises(SystemExit):
            bench.check_line(bad, 20)


def test_plates_literal_config_is_ill_conditioned(orc):
    """Config 1 as specified (two 50x50x20 EP plates on the same x-y lattice
    columns): once the weak contact lets the facing layers interpenetrate,
    particles meet head-on and the contact force direction dr/(|dr| + 1e-10)
    amplifies rounding.  Two oracle runs that differ by one ulp in the initial
    positions agree to 1e-10 in force through step 45 and differ by more than
    1e-3 at step 63, where the closest contact pair is under 1e-4 dx apart.
    The twin with the second plate shifted by half a spacing stays within 1e-10
    under the same perturbation (measured: <= 2.6e-11 in force over 500 steps,
    too long for this suite)."""
    spec = W.plates_ep()
    runs = []
    for pert in (False, True):
        e = orc.OracleEnv(spec)
        e.threaded = True
        if pert:
            e.y[:] = np.nextafter(e.y, np.inf)
        e.update_neighs()
        e.apply_bc_at0()
        snap = {}
        for i in range(1, 64):
            e.step()
            if i in (45, 63):
                snap[i] = e.f.copy()
            if i % 10 == 0:
                e.update_neighs()
        c = e.contacts[0]
        y1, y2, nb = e.y[c["mask1"]], e.y[c["mask2"]], c["neighs"]
        r, k = np.nonzero(nb > 0)
        closest = float(np.sqrt(((y1[r] - y2[nb[r, k] - 1]) ** 2).sum(axis=1)).min())
        runs.append((snap, closest))
    (a, ca), (b, _) = runs
    rel = {i: float(np.linalg.norm(b[i] - a[i]) / np.linalg.norm(a[i])) for i in (45, 63)}
    assert rel[45] <= 1e-10, rel
    assert rel[63] > 1e-3, rel
    assert ca < 1e-4, ca
