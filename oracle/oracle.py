"""ORACLE for the PeriDyn explicit-dynamics hot path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product (peridyn_amd/) never imports it.

This file restates the reference's control flow around the C++ loop restatement
in peridyn_oracle.cpp:

  * Env construction           src/environment.jl:166-235
  * GeneralMaterial            src/materials/general_material.jl:81-98
  * material constants         bond_based.jl:46-62 (c = 18K/pi/delta^4),
                               elasto_plastic.jl:59-71 (psi = 75/(8pi) sigma_e^2/delta^5),
                               shortrange.jl:23-29, utils.jl:12-25 (make_matrix)
  * force_density! latch       src/materials/general_material.jl:240-256
  * update_acc!                src/solvers/solvers.jl:320-336
  * velocity_verlet_step!      src/solvers/velocity_verlet.jl:57-89
  * BCs                        boundary_conditions.jl:94-126, FixBC.jl:46-65, ToFroBC.jl:51-111
  * contacts (RM12)            contacts.jl:76-91,189-247,403-476, nonlinearspring.jl:124-145
  * run! physics cadence       src/solvers/solvers.jl:97-187 (contact lists rebuilt once
                               before step 1 and after every step i with i % nuf == 0;
                               file writes and their update_acc! side effects omitted)
  * cal_damage                 src/peri.jl:587-594

Elementwise numpy expressions below mirror Julia broadcasts operation by
operation (numpy evaluates each binary op separately and never fuses), so the
FP64 results are the reference's.

Parity status of every function is listed in the test files; functions without a
pinning check are marked "parity unpinned" there: Julia's float `^` (contact
exponent that is not integer-valued), `volume[1]^(1/3)` for non-cube volumes, and the tie
cases of Julia's `isapprox` in the deformed latch.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

BB, OSB, EP = 0, 1, 2
_KINDS = {"bb": BB, "osb": OSB, "ep": EP}

_lib = None


def build(force: bool = False) -> str:
    """Compile the C++ restatement (g++ -O2 -ffp-contract=off, OpenMP for the
    timing-only threaded loops)."""
    src = os.path.join(_HERE, "peridyn_oracle.cpp")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.check_call(
            ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-fno-fast-math", "-fopenmp", src, "-o", _SO + ".tmp"])
        os.replace(_SO + ".tmp", _SO)
    return _SO



def _julia_int_pow(exponent) -> bool:
    """`strain^expo` with expo::Float64 (nonlinearspring.jl:16,33,140) in Julia
    1.10 (Manifest.toml: julia_version 1.10.4): `^(x::Float64, y::Float64)`
    returns `x^Int(y)` (pow_body's compensated power by squaring, restated in
    orc_julia_pow_int) when y is integer-valued, else its general power."""
    e = float(exponent)
    return math.isfinite(e) and e == math.floor(e) and abs(e) <= 4096.0

class _OrcBlock(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("Mp", ctypes.c_int64), ("ntype", ctypes.c_int64),
        ("type_start", ctypes.c_int64),
        ("fam", ctypes.c_void_p), ("intact", ctypes.c_void_p), ("x", ctypes.c_void_p),
        ("vol", ctypes.c_void_p), ("type", ctypes.c_void_p), ("s0", ctypes.c_void_p),
        ("density", ctypes.c_void_p), ("c", ctypes.c_void_p), ("K", ctypes.c_void_p),
        ("G", ctypes.c_void_p), ("psi", ctypes.c_void_p), ("m", ctypes.c_void_p),
        ("theta", ctypes.c_void_p), ("td_norm2", ctypes.c_void_p), ("edp", ctypes.c_void_p),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        P, I64, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.orc_julia_pow_int.argtypes = [D, I64]
        L.orc_julia_pow_int.restype = D
        L.orc_contact_power.argtypes = [D, D, ctypes.c_int]
        L.orc_contact_power.restype = D
        L.orc_cal_family.argtypes = [I64, P, D, I64, P, P]
        L.orc_cal_family.restype = I64
        L.orc_weighted_volume.argtypes = [I64, I64, P, P, P, P]
        L.orc_dilatation.argtypes = [I64, I64, P, P, P, P, P, P, P]
        L.orc_td_norm2.argtypes = [ctypes.POINTER(_OrcBlock), P]
        L.orc_force_density.argtypes = [ctypes.c_int, ctypes.POINTER(_OrcBlock), P, P]
        L.orc_force_density_omp.argtypes = [ctypes.c_int, ctypes.POINTER(_OrcBlock), P, P]
        L.orc_force_density_omp.restype = ctypes.c_int
        L.orc_contact_neighs.argtypes = [I64, P, I64, P, D, I64, P]
        L.orc_contact_neighs_omp.argtypes = [I64, P, I64, P, D, I64, P]
        L.orc_contact_force.argtypes = [I64, P, P, P, P, I64, P, P, P, P, I64, P, D, D, D, ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Host-side constants, restated from the reference constructors.
# ---------------------------------------------------------------------------
def julia_pow_int(x: float, n: int) -> float:
    """Julia `x::Float64 ^ n::Integer` (compensated power by squaring)."""
    return lib().orc_julia_pow_int(float(x), int(n))


def contact_power(strain: float, expo: float) -> float:
    """strain^expo as the contact force takes it (orc_contact_power)."""
    return lib().orc_contact_power(float(strain), float(expo), int(_julia_int_pow(expo)))


def make_matrix(S) -> np.ndarray:
    """src/misc/utils.jl:12-25: packed upper triangle -> symmetric N x N."""
    S = np.asarray(S, dtype=np.float64).ravel()
    n = S.size
    N = int(math.sqrt(2 * n + 0.5 ** 2) - 0.5)
    M = np.zeros((N, N))
    a = 0
    for i in range(N):
        for j in range(i, N):
            M[i, j] = S[a]
            M[j, i] = S[a]
            a += 1
    return M


def bb_bond_stiffness(K: np.ndarray, horizon: float) -> np.ndarray:
    """bond_based.jl:58 `18 * K ./ pi ./ horizon^4`."""
    h4 = julia_pow_int(horizon, 4)
    return ((18.0 * K) / math.pi) / h4


def ep_psi(sigma_y: np.ndarray, horizon: float) -> np.ndarray:
    """elasto_plastic.jl:59-71 with VonMises effective_stress2 of (sigma, 0, 0)
    (elasto_plastic.jl:364-366): sigma_e^2 = ((s-0)^2 + (0-0)^2 + (s-0)^2)/2."""
    s = np.asarray(sigma_y, dtype=np.float64)
    se2 = ((s * s + 0.0 * 0.0) + s * s) / 2.0
    return ((75.0 / (8.0 * math.pi)) * se2) / julia_pow_int(horizon, 5)


def shortrange_stiffness(K: float, horizon: float, mfactor: float = 1.0) -> float:
    """shortrange.jl:25 `18K/(pi*horizon^4) * mfactor`."""
    return ((18.0 * K) / (math.pi * julia_pow_int(horizon, 4))) * mfactor


def particle_size_of(vol: np.ndarray) -> float:
    """general_material.jl:83 `volume[1]^(1/3)` (exact for cube volumes;
    parity unpinned otherwise: Julia's float power vs C pow)."""
    v = float(vol[0])
    r = round(v ** (1.0 / 3.0))
    if r > 0 and float(r) ** 3 == v:
        return float(r)
    return math.pow(v, 1.0 / 3.0)


# ---------------------------------------------------------------------------
# Peridynamic primitives
# ---------------------------------------------------------------------------
def cal_family(x: np.ndarray, horizon: float, max_neigh: int):
    """Returns (family (n, M') int64 1-based descending zero padded, intact (n, M') uint8).
    Raises IndexError where the reference raises BoundsError."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    M = min(max_neigh, n)
    fam = np.zeros(M * n, dtype=np.int64)
    intact = np.zeros(M * n, dtype=np.uint8)
    Mp = lib().orc_cal_family(n, _p(x), float(horizon), int(max_neigh), _p(fam), _p(intact))
    if Mp == -1:
        raise IndexError("BoundsError: family degree exceeds max_neigh")
    if Mp < 0:
        raise ValueError("bad arguments")
    return fam[: Mp * n].reshape(n, Mp).copy(), intact[: Mp * n].reshape(n, Mp).copy()


def weighted_volume(x, vol, fam) -> np.ndarray:
    n, Mp = fam.shape
    m = np.zeros(n)
    lib().orc_weighted_volume(n, Mp, _p(fam), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(vol)), _p(m))
    return m


def dilatation(fam, intact, x, y, vol, m) -> np.ndarray:
    n, Mp = fam.shape
    th = np.zeros(n)
    lib().orc_dilatation(n, Mp, _p(fam), _p(intact), _p(x), _p(np.ascontiguousarray(y)), _p(vol), _p(m), _p(th))
    return th


# ---------------------------------------------------------------------------
# Self contact, ContactModel11 (src/contacts/contacts.jl:94-119, 357-380,
# 539-591), restated literally in Python (small test sizes only)
# ---------------------------------------------------------------------------
def jl_fld(x: float, y: float) -> float:
    """Julia fld for Float64: round((x - mod(x, y)) / y)."""
    r = math.fmod(x, y)
    if r != 0.0 and ((r < 0.0) != (y < 0.0)):
        r += y
    return float(round((x - r) / y))


def jl_get_cells(x: np.ndarray, horizon: float, max_part: int = 30):
    """get_cells (src/peri.jl:500-520) with neigh_cells / cell_number
    (:459-486): 1-based cell lists of 1-based particle ids, 27-stencils."""
    lo, hi = x.min(axis=0), x.max(axis=0)
    N = [int(max(1, min(max_part, jl_fld(hi[d] - lo[d], horizon)))) for d in range(3)]
    num = lambda i, j, k: i + (j - 1) * N[0] + (k - 1) * N[0] * N[1]
    cells = [[] for _ in range(N[0] * N[1] * N[2])]
    neighs = [None] * len(cells)
    for k in range(1, N[2] + 1):
        for j in range(1, N[1] + 1):
            for i in range(1, N[0] + 1):
                a = []
                for kk in (k - 1, k, k + 1):
                    for jj in (j - 1, j, j + 1):
                        for ii in (i - 1, i, i + 1):
                            if 0 < ii * jj * kk and ii <= N[0] and jj <= N[1] and kk <= N[2]:
                                a.append(num(ii, jj, kk))
                neighs[num(i, j, k) - 1] = a
    for p in range(x.shape[0]):
        c = [int(max(1, min(max_part, jl_fld(x[p, d] - lo[d], horizon)))) for d in range(3)]
        cells[num(*c) - 1].append(p + 1)
    return cells, neighs


def rm11_neighs(x, fam, intact, search_distance, equi_dist, max_neighs, max_part=30) -> np.ndarray:
    """update_contact_neighs! for ContactModel11 (contacts.jl:539-591) as
    written: particles with a broken bond, cells of width search_distance,
    and for every stencil cell a merge walk of its sorted ids against the
    particle's family column, which the walk expects ascending but which
    GeneralMaterial stores descending (general_material.jl:88-90).  fam, intact:
    (n, M') rows.  Returns neighbours as (n, max_neighs), 1-based, 0-padded."""
    n = x.shape[0]
    FS = fam.shape[1]
    mask = np.any((fam > 0) & (intact == 0), axis=1)
    cells, cell_neighs = jl_get_cells(x, search_distance, max_part)
    out = np.zeros((n, max_neighs), dtype=np.int64)
    e2 = equi_dist * equi_dist
    for cell_i in range(len(cells)):
        for ca in cells[cell_i]:
            ind = 1
            if not mask[ca - 1]:
                continue
            a1, b1, c1 = x[ca - 1]
            fms = fam[ca - 1]
            for nc in cell_neighs[cell_i]:
                nbs = cells[nc - 1]
                nbs.sort()
                fm_ind = 1
                for neigh in nbs:
                    while fms[fm_ind - 1] < neigh and fm_ind < FS:
                        fm_ind += 1
                    if fms[fm_ind - 1] == neigh and not intact[ca - 1, fm_ind - 1]:
                        a2, b2, c2 = x[neigh - 1]
                        dr2 = ((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2)) + (c1 - c2) * (c1 - c2)
                        if dr2 < e2:
                            out[ca - 1, ind - 1] = neigh
                            ind += 1
    return out


def rm11_force(y, f, vol, den, neighs, equi, K, expo, expo_is_int):
    """short_range_contact! for ContactModel11 (contacts.jl:357-380): i in
    order, f_i += F (vol_j / den_i), f_j -= F (vol_j / den_j) (vol_j on both
    sides, as written), F = NonLinearSpring (nonlinearspring.jl:124-145) of
    dr = y_i - y_j."""
    n, M = neighs.shape
    for i in range(n):
        for k in range(M):
            j = int(neighs[i, k])
            if j == 0:
                continue
            j -= 1
            d = y[i] - y[j]
            mag = math.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            dl = equi - mag
            if dl < 0.0:
                F = np.zeros(3)
            else:
                st = dl / equi
                pw = lib().orc_julia_pow_int(st, int(expo)) if expo_is_int else st ** expo
                F = ((K * pw) / (mag + 1e-10)) * d
            f[i] += F * (vol[j] / den[i])
            f[j] -= F * (vol[j] / den[j])


# ---------------------------------------------------------------------------
# Blocks and environment
# ---------------------------------------------------------------------------
class OracleBlock:
    def __init__(self, b: dict):
        self.kind = _KINDS[b["kind"]]
        self.x = np.ascontiguousarray(b["x"], dtype=np.float64)
        self.y0 = np.ascontiguousarray(b.get("y", b["x"]), dtype=np.float64)
        self.v0 = np.ascontiguousarray(b.get("v", np.zeros_like(self.x)), dtype=np.float64)
        self.vol = np.ascontiguousarray(b["vol"], dtype=np.float64)
        self.type = np.ascontiguousarray(b.get("type", np.ones(self.x.shape[0], np.int64)), dtype=np.int64)
        self.n = self.x.shape[0]
        self.horizon = float(b["horizon"])
        self.type_start = int(self.type.min())
        self.ntype = int(self.type.max()) - self.type_start + 1
        self.particle_size = float(b["particle_size"]) if b.get("particle_size") else particle_size_of(self.vol)
        if "family" in b:
            self.fam = np.ascontiguousarray(b["family"], dtype=np.int64)
            self.intact = np.ascontiguousarray(b["intact"], dtype=np.uint8).copy()
        else:
            self.fam, self.intact = cal_family(self.x, self.horizon, int(b.get("max_neigh", 100)))
        self.Mp = self.fam.shape[1]
        self.m = weighted_volume(self.x, self.vol, self.fam)
        self.deformed = not np.array_equal(self.y0, self.x)
        nt = self.ntype

        def mat(key, default=None):
            v = b.get(key, default)
            a = np.asarray(v, dtype=np.float64)
            if a.ndim < 2:
                a = make_matrix(a.ravel())
            assert a.shape == (nt, nt), (key, a.shape, nt)
            return np.asfortranarray(a).ravel(order="F").copy()

        self.density = np.ascontiguousarray(np.asarray(b["density"], dtype=np.float64).ravel())
        self.s0 = mat("critical_stretch")
        self.c = np.zeros(nt * nt)
        self.K = np.zeros(nt * nt)
        self.G = np.zeros(nt * nt)
        self.psi = np.zeros(nt * nt)
        if self.kind == BB:
            Km = np.asarray(b["bulk_modulus"], dtype=np.float64)
            Km = make_matrix(Km.ravel()) if Km.ndim < 2 else Km
            self.c = np.asfortranarray(bb_bond_stiffness(Km, self.horizon)).ravel(order="F").copy()
        else:
            self.K = mat("bulk_modulus")
            self.G = mat("shear_modulus")
            if self.kind == EP:
                sy = np.asarray(b["yield_stress"], dtype=np.float64)
                sy = make_matrix(sy.ravel()) if sy.ndim < 2 else sy
                self.psi = np.asfortranarray(ep_psi(sy, self.horizon)).ravel(order="F").copy()
        self.theta = np.zeros(self.n)
        self.td_norm2 = np.zeros(self.n)
        self.edp = np.zeros((self.n, self.Mp))
        self._s = self._struct()

    def _struct(self) -> _OrcBlock:
        s = _OrcBlock()
        s.n, s.Mp, s.ntype, s.type_start = self.n, self.Mp, self.ntype, self.type_start
        for name in ("fam", "intact", "x", "vol", "type", "s0", "density", "c", "K", "G", "psi",
                     "m", "theta", "td_norm2", "edp"):
            setattr(s, name, _p(getattr(self, name)).value)
        return s

    def force_density(self, y_blk: np.ndarray, f_blk: np.ndarray, threaded: bool = False) -> int:
        """Adds the block's force density to f_blk; returns the threads used.
        threaded: OpenMP loops, for timing only (f then differs from the
        single-threaded sums by rounding)."""
        if threaded:
            return lib().orc_force_density_omp(self.kind, ctypes.byref(self._s), _p(y_blk), _p(f_blk))
        lib().orc_force_density(self.kind, ctypes.byref(self._s), _p(y_blk), _p(f_blk))
        return 1


class OracleEnv:
    """Env + velocity-Verlet run loop of the reference, restated on the CPU."""

    def __init__(self, spec: dict):
        self.dt = float(spec["dt"])
        self.blocks = [OracleBlock(b) for b in spec["blocks"]]
        self.y = np.ascontiguousarray(np.concatenate([b.y0 for b in self.blocks]))
        self.v = np.ascontiguousarray(np.concatenate([b.v0 for b in self.blocks]))
        self.f = np.zeros_like(self.y)
        self.N = self.y.shape[0]
        self.volume = np.concatenate([b.vol for b in self.blocks])
        self.type = np.concatenate([b.type for b in self.blocks])
        mass = []
        self.limits = []
        lo = 0
        for b in self.blocks:
            mm = b.vol.copy()
            for j in range(b.type_start, b.type_start + b.ntype):
                mk = b.type == j
                mm[mk] = mm[mk] * b.density[j - b.type_start]
            mass.append(mm)
            self.limits.append((lo, lo + b.n))
            lo += b.n
        self.mass = np.concatenate(mass)
        self.density = self.mass / self.volume
        self.intact0 = np.concatenate([b.intact.sum(axis=1) for b in self.blocks]).astype(np.int64)
        self.contacts = []
        for c in spec.get("contacts", []):
            self.contacts.append(self._make_contact(c))
        self.bcs = []
        for bc in spec.get("bcs", []):
            d = dict(bc)
            d["mask"] = np.asarray(bc["mask"], dtype=bool)
            d["last"] = np.zeros((int(d["mask"].sum()), 3))
            d["dir"] = 1
            d["onlyatstart"] = bool(bc.get("onlyatstart", False))
            dims = bc.get("dims")
            if dims is not None and not all(bool(v) for v in dims):
                # y[dims, bool] .= (3 x n update) (boundary_conditions.jl:108-110)
                raise ValueError("DimensionMismatch: a BC on a subset of dims cannot take its 3 x n update")
            if d["kind"] == "tofro":
                d["rate"] = np.asarray(bc["rate"], dtype=np.float64)
                d.setdefault("freq", math.inf)
                d.setdefault("applyafter", 0)
            self.bcs.append(d)
        self.time_step = 0
        # threaded: the reference's threaded loops on OpenMP threads, for the
        # CPU baseline's timing only; threads_used is what the last force pass used
        self.threaded = False
        self.threads_used = 1

    # -- contacts ----------------------------------------------------------
    def _make_contact(self, c: dict) -> dict:
        if len(c["pair"]) == 1 or len(c["pair"][1]) == 0:
            return self._make_self_contact(c)
        t1, t2 = [set(int(t) for t in ts) for ts in c["pair"]]
        mask1 = np.isin(self.type, list(t1))
        mask2 = np.isin(self.type, list(t2))
        assert not np.any(mask1 & mask2), "contact pair type sets must be disjoint"
        b1 = next(b for b in self.blocks if b.type_start in t1)
        b2 = next(b for b in self.blocks if b.type_start in t2)
        distanceD = float(c.get("distanceD", 1.0))
        distanceX = float(c.get("distanceX", 3.0))
        # ContactModel12_gcal (contacts.jl:76-91)
        equi = (b1.particle_size + b2.particle_size) / 2.0 * distanceD
        n2 = int(mask2.sum())
        max_neighs = min(int(c.get("max_neighs", 50)), n2)
        expo = c.get("exponent", 1)
        return dict(mask1=mask1, mask2=mask2, equi=equi, distance=distanceX * equi,
                    max_neighs=max_neighs, K=float(c["stiffness"]),
                    expo=float(expo), expo_is_int=int(_julia_int_pow(expo)),
                    neighs=np.zeros((int(mask1.sum()), max_neighs), dtype=np.int64))

    def _make_self_contact(self, c: dict) -> dict:
        """ContactModel11_gcal (contacts.jl:105-119): one block, equi_dist =
        particle_size * distanceD, distance = distanceX * equi_dist."""
        t1 = set(int(t) for t in c["pair"][0])
        bi = next(i for i, b in enumerate(self.blocks) if b.type_start in t1)
        b = self.blocks[bi]
        equi = b.particle_size * float(c.get("distanceD", 1.0))
        max_neighs = min(int(c.get("max_neighs", 50)), b.n)
        expo = c.get("exponent", 1)
        return dict(self_block=bi, equi=equi, distance=float(c.get("distanceX", 5.0)) * equi,
                    max_neighs=max_neighs, K=float(c["stiffness"]), expo=float(expo),
                    expo_is_int=int(_julia_int_pow(expo)),
                    neighs=np.zeros((b.n, max_neighs), dtype=np.int64))

    def update_neighs(self):
        """update_neighs! (solvers.jl:351-361) -> update_contact_neighs!."""
        for c in self.contacts:
            if "self_block" in c:
                b = self.blocks[c["self_block"]]
                lo, hi = self.limits[c["self_block"]]
                c["neighs"] = rm11_neighs(self.y[lo:hi], b.fam, b.intact, c["distance"], c["equi"],
                                          c["max_neighs"])
                continue
            y1 = np.ascontiguousarray(self.y[c["mask1"]])
            y2 = np.ascontiguousarray(self.y[c["mask2"]])
            search = lib().orc_contact_neighs_omp if self.threaded else lib().orc_contact_neighs
            search(y1.shape[0], _p(y1), y2.shape[0], _p(y2), c["distance"], c["max_neighs"], _p(c["neighs"]))

    def _contact_force(self, c: dict):
        if "self_block" in c:
            lo, hi = self.limits[c["self_block"]]
            f = self.f[lo:hi]
            rm11_force(self.y[lo:hi], f, self.volume[lo:hi], self.density[lo:hi], c["neighs"], c["equi"],
                       c["K"], c["expo"], c["expo_is_int"])
            return
        m1, m2 = c["mask1"], c["mask2"]
        y1 = np.ascontiguousarray(self.y[m1]); y2 = np.ascontiguousarray(self.y[m2])
        f1 = np.ascontiguousarray(self.f[m1]); f2 = np.ascontiguousarray(self.f[m2])
        v1 = np.ascontiguousarray(self.volume[m1]); v2 = np.ascontiguousarray(self.volume[m2])
        d1 = np.ascontiguousarray(self.density[m1]); d2 = np.ascontiguousarray(self.density[m2])
        lib().orc_contact_force(y1.shape[0], _p(y1), _p(v1), _p(d1), _p(f1),
                                y2.shape[0], _p(y2), _p(v2), _p(d2), _p(f2),
                                c["max_neighs"], _p(c["neighs"]), c["equi"], c["K"],
                                c["expo"], c["expo_is_int"])
        self.f[m1] = f1
        self.f[m2] = f2

    # -- forces ------------------------------------------------------------
    def _undeformed(self, b: OracleBlock, lo: int, hi: int) -> bool:
        """isapprox(y_ .- y_[:,1] .+ x[:,1], x) with rtol = sqrt(eps())."""
        yb = self.y[lo:hi]
        a = (yb - yb[0]) + b.x[0]
        d = float(np.linalg.norm((a - b.x).ravel()))
        if math.isfinite(d):
            return d <= math.sqrt(np.finfo(np.float64).eps) * max(
                float(np.linalg.norm(a.ravel())), float(np.linalg.norm(b.x.ravel())))
        return bool(np.all(a == b.x))

    def update_acc(self):
        self.f.fill(0.0)
        for b, (lo, hi) in zip(self.blocks, self.limits):
            if not b.deformed:
                if self._undeformed(b, lo, hi):
                    continue
                b.deformed = True
            self.threads_used = b.force_density(self.y[lo:hi], self.f[lo:hi], self.threaded)
        for c in self.contacts:
            self._contact_force(c)
        # update_misc! (solvers.jl:289-291): momentum from the current velocities
        self.p = self.mass[:, None] * self.v

    def printdata(self) -> np.ndarray:
        """printdata columns (environment.jl:308-320): time, sum(p), sum(f),
        mean(cal_damage).  Sums in numpy's pairwise order, as Julia's sum."""
        p = getattr(self, "p", np.zeros_like(self.v))
        d = self.damage()
        return np.array([self.time_step * self.dt, *p.sum(axis=0), *self.f.sum(axis=0), d.sum() / d.size])

    # -- boundary conditions -------------------------------------------------
    def apply_bc_at0(self):
        for bc in self.bcs:
            bc["last"] = self.y[bc["mask"]].copy()

    def _bc_position(self):
        for bc in self.bcs:
            if bc["kind"] == "fix":
                self.y[bc["mask"]] = bc["last"]
            else:
                ynew = bc["last"] + (self.dt * float(bc["dir"])) * bc["rate"]
                self.y[bc["mask"]] = ynew
                bc["last"] = ynew

    def _bc_velocity(self):
        for bc in self.bcs:
            if bc["kind"] == "fix":
                self.v[bc["mask"]] = 0.0
            else:
                self.v[bc["mask"]] = float(bc["dir"]) * bc["rate"]

    def _bc_check(self):
        for bc in self.bcs:
            if bc["kind"] == "tofro":
                t = self.time_step
                fr = bc["freq"]
                rem = math.fmod(float(t), float(fr)) if isinstance(fr, float) else t % fr
                if t > bc["applyafter"] and rem == 0:
                    bc["dir"] = -bc["dir"]

    # -- integrator ----------------------------------------------------------
    def step(self):
        """velocity_verlet_step! (velocity_verlet.jl:57-89)."""
        c = self.dt / 2.0
        self.v += c * self.f
        self.y += self.dt * self.v
        self._bc_position()
        self.update_acc()
        self.v += c * self.f
        self._bc_velocity()
        self._bc_check()
        self.time_step += 1

    def run(self, nsteps: int, nuf: int = 1, start: bool = True, log_every: int = 0):
        """run! physics cadence (solvers.jl:97-187) without file output; with
        log_every > 0 returns the printdata rows of step 1 and of every
        log_every-th step (solvers.jl:158-167), taken before that step's
        neighbour update."""
        if start:
            self.update_neighs()
            self.apply_bc_at0()
        rows = []
        for i in range(1, nsteps + 1):
            self.step()
            if log_every > 0 and (i % log_every == 0 or i == 1):
                rows.append(self.printdata())
            if i % nuf == 0:
                self.update_neighs()
        return np.array(rows) if log_every > 0 else None

    def minimize(self, step_size: float, drag: float, max_iter: int = 1000, x_tol: float = 1e-3,
                 f_tol: float = 1e-3):
        """QSDrag minimize! (minimize.jl:105-231): BCs once, then update_acc!
        and the drag update of the free particles until max|f| < f_tol and
        max|dy| < x_tol; the clipped velocity of apply_! is not stored back
        (it rebinds a local), so opt keeps its unclipped value.  Returns
        (iterations, max|f|, max|dy|) of the last iteration."""
        self._bc_position()
        self._bc_velocity()
        dt = step_size
        a = self.blocks[0].particle_size / 10.0
        held = np.zeros(self.N, dtype=bool)
        for bc in self.bcs:
            if not bc["onlyatstart"]:  # minimize.jl:138-143
                held |= bc["mask"]
        free = ~held
        opt = np.zeros((int(free.sum()), 3))
        k = drag / 10.0
        it, fm, xm = 0, math.nan, math.nan
        for i in range(1, max_iter + 1):
            self.update_acc()
            gs = self.f[free]
            force = gs - (drag * opt) * (1.0 + k * np.abs(opt))
            dy = (0.5 * force) * (dt * dt) + opt * dt
            opt = opt + dt * force
            dy = np.maximum(np.minimum(dy, a), -a)
            self.y[free] += dy
            it, fm, xm = i, float(np.abs(gs).max()), float(np.abs(dy).max())
            self.qs_opt = opt
            if fm < f_tol and xm < x_tol:
                break
        self._bc_check()
        self.time_step += 1
        return it, fm, xm

    def damage(self) -> np.ndarray:
        counts = np.concatenate([b.intact.sum(axis=1) for b in self.blocks]).astype(np.float64)
        return 1.0 - counts / self.intact0.astype(np.float64)

    def intact_all(self) -> list:
        return [b.intact.copy() for b in self.blocks]
