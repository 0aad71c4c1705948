"""Synthetic specimens and the benchmark workloads (plain data "specs").

The reference builds geometry with PDMaterialPoints (`create(Cuboid(...))`,
examples/tensile/bar.jl:24-25, examples/collision/impact_disk.jl:16-17,43-44),
an un-vendored dependency.  Here every specimen is an integer lattice of spacing
dx with cell volume dx^3, generated in x-fastest order (index = i + nx*(j + ny*k)),
which keeps neighbours close in memory.

A spec is a plain dict consumed both by the HIP path (peridyn_amd.Env.from_spec)
and by the CPU oracle used in tests:

  {"dt": float,
   "blocks":  [{"kind": "bb"|"osb"|"ep", "x": (n,3), "y": (n,3), "v": (n,3),
                "vol": (n,), "type": (n,) int64, "horizon": h, "max_neigh": M,
                "bulk_modulus": K, "shear_modulus": G, "critical_stretch": s0,
                "density": rho, "yield_stress": sy}],
   "contacts": [{"pair": (types1, types2), "exponent": p, "stiffness": k,
                 "distanceD": dD, "distanceX": dX, "max_neighs": m}],
   "bcs":      [{"kind": "fix", "mask": (N,) bool} |
                {"kind": "tofro", "mask": (N,) bool, "rate": (3,), "freq": f, "applyafter": a}]}

Material parameters follow the reference constructors: scalars or packed upper
triangles (make_matrix, src/misc/utils.jl:12-25) or full ntype x ntype matrices.
"""
from __future__ import annotations

import math

import numpy as np


def lattice_box(nx: int, ny: int, nz: int, dx: float = 1.0, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    x = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1) * dx
    return np.ascontiguousarray(x + np.asarray(origin, dtype=np.float64))


def lattice_disk(radius: float, nz: int, dx: float = 1.0) -> np.ndarray:
    """Lattice points with x^2 + y^2 <= radius^2 (in dx units), z in [0, nz)."""
    r = int(math.floor(radius))
    ax = np.arange(-r, r + 1, dtype=np.float64)
    Y, X = np.meshgrid(ax, ax, indexing="ij")
    keep = (X * X + Y * Y) <= radius * radius
    xy = np.stack([X[keep], Y[keep]], axis=1)
    layers = [np.concatenate([xy, np.full((xy.shape[0], 1), float(z))], axis=1) for z in range(nz)]
    return np.ascontiguousarray(np.concatenate(layers) * dx)


def lattice_sphere(radius: float, center=(0.0, 0.0, 0.0), dx: float = 1.0) -> np.ndarray:
    r = int(math.floor(radius))
    ax = np.arange(-r, r + 1, dtype=np.float64)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    keep = (X * X + Y * Y + Z * Z) <= radius * radius
    x = np.stack([X[keep], Y[keep], Z[keep]], axis=1) * dx
    return np.ascontiguousarray(x + np.asarray(center, dtype=np.float64))


def bulk_from_E(E: float, nu: float = 0.25) -> float:
    return E / (3.0 * (1.0 - 2.0 * nu))


def block(kind: str, x: np.ndarray, horizon: float, *, type_id: int = 1, v=None, y=None,
          max_neigh: int = 200, **params) -> dict:
    n = x.shape[0]
    b = {"kind": kind, "x": x, "y": x.copy() if y is None else y,
         "v": np.zeros_like(x) if v is None else np.broadcast_to(np.asarray(v, dtype=np.float64), x.shape).copy(),
         "vol": np.ones(n), "type": np.full(n, type_id, dtype=np.int64),
         "horizon": float(horizon), "max_neigh": int(max_neigh)}
    b.update(params)
    return b


def grips_z(x: np.ndarray, layers: int, rate: float) -> list:
    """MoveBC grips on the outermost `layers` z-layers pulled apart at +-rate
    (rate is a velocity, ToFroBC.jl:56-60)."""
    z = x[:, 2]
    lo, hi = z.min(), z.max()
    bot = z <= lo + (layers - 1) + 1e-9
    top = z >= hi - (layers - 1) - 1e-9
    return [{"kind": "tofro", "mask": bot, "rate": np.array([0.0, 0.0, -rate]), "freq": math.inf, "applyafter": 0},
            {"kind": "tofro", "mask": top, "rate": np.array([0.0, 0.0, rate]), "freq": math.inf, "applyafter": 0}]


# ---------------------------------------------------------------------------
# Benchmark workloads.  Parameters the workload descriptions leave open are fixed
# here once (and mirrored in fixtures/configs/*.json).
# ---------------------------------------------------------------------------
E0, NU0, RHO0, S00, DELTA = 70e3, 0.25, 2.7e-3, 0.005, 3.015


def tensile_bb(n: int = 20, dt: float = 1e-3, pull_per_step: float = 0.01, kind: str = "bb",
               s0: float = S00, E: float = E0, layers: int = 3) -> dict:
    """Tensile block: n^3 lattice, grips on 3 z-layers each end pulled at
    +-pull_per_step dx per step (rate = pull_per_step/dt)."""
    x = lattice_box(n, n, n)
    K = bulk_from_E(E, NU0)
    params = dict(bulk_modulus=K, critical_stretch=s0, density=RHO0)
    if kind != "bb":
        params["shear_modulus"] = 3.0 * K * (1 - 2 * NU0) / (2 * (1 + NU0))
    if kind == "ep":
        params["yield_stress"] = 100.0
    b = block(kind, x, DELTA, **params)
    return {"dt": dt, "blocks": [b], "contacts": [], "bcs": grips_z(x, layers, pull_per_step / dt)}


def plates_ep(nx: int = 50, ny: int = 50, nz: int = 20, gap: float = 1.0, dt: float = 5e-5,
              approach_per_step: float = 0.05, exponent: int = 2, shift: float = 0.0) -> dict:
    """Two elastoplastic plates approaching each other (state-based collision).

    shift: lateral offset (dx units, in x and y) of the second plate.  With 0
    (the workload as specified) both plates share their x-y columns, and once
    the weak contact lets the facing layers interpenetrate (from about step 47)
    particles meet head-on at distances down to 1e-6 dx, where the contact
    force's direction dr/(|dr| + 1e-10) amplifies any rounding difference about
    5x per step.  shift = 0.5 puts the second plate's columns between the
    first's, so facing particles never come closer than 0.5*sqrt(2) dx
    laterally: the well-conditioned twin used for long FP64 comparisons."""
    xa = lattice_box(nx, ny, nz)
    xb = lattice_box(nx, ny, nz, origin=(float(shift), float(shift), float(nz - 1) + 1.0 + gap))
    vz = 0.5 * approach_per_step / dt
    p = dict(bulk_modulus=50e3, shear_modulus=25e3, critical_stretch=0.01, density=RHO0, yield_stress=100.0)
    ba = block("ep", xa, DELTA, type_id=1, v=(0.0, 0.0, vz), **p)
    bb = block("ep", xb, DELTA, type_id=2, v=(0.0, 0.0, -vz), **p)
    K_contact = 18.0 * 50e3 / (math.pi * DELTA ** 4)
    contact = {"pair": ([1], [2]), "exponent": exponent, "stiffness": K_contact,
               "distanceD": 0.9, "distanceX": 2.0, "max_neighs": 50}
    return {"dt": dt, "blocks": [ba, bb], "contacts": [contact], "bcs": []}


def disk_impact(radius: float = 150.0, thickness: int = 10, proj_radius: float = 15.0,
                dt: float = 1e-5, speed_per_step: float = 0.2, gap: float = 2.0) -> dict:
    """Bond-based disk hit by a stiff spherical projectile (linear contact)."""
    xd = lattice_disk(radius, thickness)
    top = float(thickness - 1)
    xs = lattice_sphere(proj_radius, center=(0.0, 0.0, top + gap + proj_radius))
    K = bulk_from_E(E0, NU0)
    bd = block("bb", xd, DELTA, type_id=1, bulk_modulus=K, critical_stretch=0.003, density=RHO0)
    bs = block("bb", xs, DELTA, type_id=2, v=(0.0, 0.0, -speed_per_step / dt),
               bulk_modulus=100.0 * K, critical_stretch=1.0, density=RHO0)
    K_contact = 18.0 * K / (math.pi * DELTA ** 4)
    contact = {"pair": ([2], [1]), "exponent": 1, "stiffness": K_contact,
               "distanceD": 1.0, "distanceX": 2.0, "max_neighs": 50}
    return {"dt": dt, "blocks": [bd, bs], "contacts": [contact], "bcs": []}


def batched_specimens(count: int = 256, seed: int = 20261203, nmin: int = 2000, nmax: int = 60000,
                      dt: float = 1e-4) -> list:
    """Independent tensile specimens: particle counts log-uniform in [nmin, nmax]
    as near-cubic lattice boxes, E and s0 uniform in [0.5x, 2x] of the tensile
    block values, pull rates uniform in [0.005, 0.02] dx/step, material 50/50
    bond-based / ordinary state-based (elastic)."""
    rng = np.random.default_rng(seed)
    specs = []
    for s in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        side = max(4, int(round(n ** (1.0 / 3.0))))
        nz = max(7, int(round(n / (side * side))))
        E = E0 * rng.uniform(0.5, 2.0)
        s0 = S00 * rng.uniform(0.5, 2.0)
        pull = rng.uniform(0.005, 0.02)
        kind = "bb" if (s % 2 == 0) else "osb"
        x = lattice_box(side, side, nz)
        K = bulk_from_E(E, NU0)
        params = dict(bulk_modulus=K, critical_stretch=s0, density=RHO0)
        if kind == "osb":
            params["shear_modulus"] = 3.0 * K * (1 - 2 * NU0) / (2 * (1 + NU0))
        b = block(kind, x, DELTA, **params)
        specs.append({"dt": dt, "blocks": [b], "contacts": [], "bcs": grips_z(x, 3, pull / dt)})
    return specs


def lattice_ball_count(horizon: float) -> int:
    """Number of unit-lattice offsets xi != 0 with |xi| < horizon: the degree of
    an interior particle of a unit lattice, the maximum degree of any lattice
    specimen with that horizon (cal_family!'s strict test, src/peri.jl:545)."""
    r = int(math.ceil(horizon))
    ax = np.arange(-r, r + 1)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    d2 = X * X + Y * Y + Z * Z
    return int(np.count_nonzero((d2 > 0) & (d2 < horizon * horizon)))


def indentation(n: int = 20000, horizon: float = DELTA, kind: str = "bb", steps: int = 1000,
                dt: float = 1e-5, E: float = E0, s0: float = 0.1) -> dict:
    """Block indentation in the shape of examples/indentation/block_indentor.jl:22-99:
    a 2a x 2a x a block (about n particles) held by FixBC on its 3 bottom
    layers, a sphere of radius a/4 placed 2 dx above it and driven down by
    ToFroBC at 10 dx over the run, reversing after 80% of the steps; 100x
    stiffer bond-based indenter; linear contact k = 18K/(pi delta^4) with the
    contact model's default distances (distanceD 1, distanceX 3, max_neighs 50,
    src/contacts/contacts.jl:76-91).  The example's self-contact models (RM1,
    RM2) are not part of this workload."""
    a = max(8, int(round((n / 4.0) ** (1.0 / 3.0))))
    xb = lattice_box(2 * a, 2 * a, a)
    top = float(a - 1)
    rs = max(2.0, a / 4.0)
    xs = lattice_sphere(rs, center=(a - 0.5, a - 0.5, top + 2.0 + rs))
    K = bulk_from_E(E, NU0)
    maxn = lattice_ball_count(horizon) + 8
    p = dict(bulk_modulus=K, critical_stretch=s0, density=RHO0)
    if kind != "bb":
        p["shear_modulus"] = 3.0 * K * (1 - 2 * NU0) / (2 * (1 + NU0))
    bb = block(kind, xb, horizon, type_id=1, max_neigh=maxn, **p)
    bs = block("bb", xs, horizon, type_id=2, max_neigh=maxn, bulk_modulus=100.0 * K, critical_stretch=1.0,
               density=RHO0)
    N = xb.shape[0] + xs.shape[0]
    fix = np.zeros(N, dtype=bool)
    fix[: xb.shape[0]] = xb[:, 2] <= 2.0 + 1e-9
    ind = np.zeros(N, dtype=bool)
    ind[xb.shape[0]:] = True
    speed = 10.0 / (steps * dt)
    bcs = [{"kind": "fix", "mask": fix},
           {"kind": "tofro", "mask": ind, "rate": np.array([0.0, 0.0, -speed]), "freq": float(int(steps * 0.8)),
            "applyafter": 0}]
    contact = {"pair": ([2], [1]), "exponent": 1, "stiffness": 18.0 * K / (math.pi * horizon ** 4),
               "distanceD": 1.0, "distanceX": 3.0, "max_neighs": 50}
    return {"dt": dt, "blocks": [bb, bs], "contacts": [contact], "bcs": bcs}


def tensile(n: int, horizon: float = DELTA, kind: str = "bb", pull_per_step: float = 0.01, dt: float = 1e-4,
            E: float = E0, s0: float = S00) -> dict:
    """Near-cubic tensile box of about n particles (>= 7 z-layers), MoveBC grips
    on 3 z-layers each end."""
    side = max(4, int(round(n ** (1.0 / 3.0))))
    nz = max(7, int(round(n / (side * side))))
    x = lattice_box(side, side, nz)
    K = bulk_from_E(E, NU0)
    p = dict(bulk_modulus=K, critical_stretch=s0, density=RHO0)
    if kind != "bb":
        p["shear_modulus"] = 3.0 * K * (1 - 2 * NU0) / (2 * (1 + NU0))
    b = block(kind, x, horizon, max_neigh=lattice_ball_count(horizon) + 8, **p)
    return {"dt": dt, "blocks": [b], "contacts": [], "bcs": grips_z(x, 3, pull_per_step / dt)}


def served_requests(count: int = 64, seed: int = 20261204, nmin: int = 1000, nmax: int = 200000,
                    smin: int = 200, smax: int = 2000, rate: float = 10.0, step_quantum: int = 10) -> list:
    """Request stream: Poisson arrivals at `rate` per second; particle counts
    log-uniform in [nmin, nmax]; step counts uniform in [smin, smax] rounded to
    multiples of `step_quantum` (the contact-rebuild period, so that a request
    stepped in chunks rebuilds at the same steps as alone); horizon uniform in
    [2.5, 4.0] dx; material bond-based / ordinary state-based 50/50; tensile or
    indentation 50/50.  Returns dicts {id, arrival, steps, nuf, kind, shape,
    n, horizon, spec_args}; make_spec(r) builds the spec."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(round(rng.uniform(smin, smax) / step_quantum)) * step_quantum
        h = float(rng.uniform(2.5, 4.0))
        kind = "bb" if rng.uniform() < 0.5 else "osb"
        shape = "tensile" if rng.uniform() < 0.5 else "indentation"
        out.append({"id": i, "arrival": t, "steps": steps, "nuf": step_quantum, "kind": kind, "shape": shape,
                    "n": n, "horizon": h})
    return out


def make_spec(r: dict) -> dict:
    if r["shape"] == "tensile":
        return tensile(r["n"], horizon=r["horizon"], kind=r["kind"])
    return indentation(r["n"], horizon=r["horizon"], kind=r["kind"], steps=r["steps"])
