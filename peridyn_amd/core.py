"""Resident simulation environments on one MI355X, driven through the C ABI.

`Env.from_spec(spec)` uploads a spec (see peridyn_amd.workloads) and builds every
block's neighbour family on the GPU; `Env.run(nsteps, nuf)` steps it with the
reference's run! cadence; `Batch` steps many envs together in one launch per
kernel class.  Arrays cross the boundary in the reference's layouts: positions
(n, 3) C-order (= Julia 3 x n column-major), family (n, rows) (= M' x n).
"""
from __future__ import annotations

import ctypes
import math
from typing import Iterable, List, Optional

import numpy as np

from . import _lib
from . import materials as M

_KIND = {"bb": _lib.PD_BB, "osb": _lib.PD_OSB, "ep": _lib.PD_EP}
_PREC = {"exact": _lib.PD_PREC_EXACT, "fp64": _lib.PD_PREC_FP64, "fp32c": _lib.PD_PREC_FP32C}


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)



def pow_by_squaring(exponent) -> bool:
    """Whether the reference evaluates `strain^exponent` by power by squaring.
    The contact models store the exponent as Float64 (nonlinearspring.jl:16,33),
    and Julia 1.10's `^(::Float64, ::Float64)` (the reference's Manifest pins
    1.10.4) returns `x^Int(y)`, i.e. pow_body's compensated squaring, whenever y
    is integer-valued; Int exponents are converted to Float64 on construction.
    Exponents beyond +-4096 are sent to the general power."""
    e = float(exponent)
    return math.isfinite(e) and e == math.floor(e) and abs(e) <= 4096.0

class Context:
    """One HIP device + stream (pd_ctx)."""

    _default = None

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.pd_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def check(self, rc: int):
        _lib.check(rc, self.h)

    def sync(self):
        self.check(self.lib.pd_sync(self.h))

    def mem_info(self):
        """(free, total) device bytes of the context's GPU."""
        fr, tot = ctypes.c_int64(0), ctypes.c_int64(0)
        self.check(self.lib.pd_ctx_mem_info(self.h, ctypes.byref(fr), ctypes.byref(tot)))
        return fr.value, tot.value

    def profile(self, enable: bool = True):
        """Start (and reset) or stop per-kernel-class HIP-event timing."""
        self.check(self.lib.pd_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self, cls: str):
        """(total device ms, launches) of a kernel class since profile(True)."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self.check(self.lib.pd_ctx_profile_read(self.h, _lib.PROF_CLASSES[cls], ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            self.lib.pd_ctx_destroy(self.h)
            self.h = None
            if Context._default is self:
                Context._default = None


class Env:
    def __init__(self, ctx: Context, env_id: int, nblocks: int, N: int, block_first: List[int],
                 block_n: List[int], has_ep: bool, spec: dict):
        self.ctx, self.id, self.nblocks, self.N = ctx, env_id, nblocks, N
        self.block_first, self.block_n, self.has_ep = block_first, block_n, has_ep
        self.spec = spec

    # -- construction ---------------------------------------------------------
    @classmethod
    def from_spec(cls, spec: dict, precision: str = "fp64", ctx: Optional[Context] = None,
                  lanes_per_particle: int = 0) -> "Env":
        ctx = ctx or Context.default()
        L = ctx.lib
        eid = ctypes.c_int64()
        ctx.check(L.pd_env_create(ctx.h, float(spec["dt"]), _PREC[precision], ctypes.byref(eid)))
        env_id = eid.value
        if lanes_per_particle:
            ctx.check(L.pd_env_set_schedule(ctx.h, env_id, int(lanes_per_particle)))
        keep = []
        firsts, ns = [], []
        N = 0
        has_ep = False
        psizes = []
        type_ranges = []
        try:
            for b in spec["blocks"]:
                x = np.ascontiguousarray(b["x"], dtype=np.float64)
                n = x.shape[0]
                y = np.ascontiguousarray(b.get("y", x), dtype=np.float64)
                v = np.ascontiguousarray(b.get("v", np.zeros_like(x)), dtype=np.float64)
                vol = np.ascontiguousarray(b["vol"], dtype=np.float64)
                typ = np.ascontiguousarray(b.get("type", np.ones(n, np.int64)), dtype=np.int64)
                t0, t1 = int(typ.min()), int(typ.max())
                nt = t1 - t0 + 1
                kind = b["kind"]
                dens = np.ascontiguousarray(np.asarray(b["density"], dtype=np.float64).ravel())
                if dens.size != nt:
                    raise ValueError("density needs one value per type")

                def fmat(a):
                    m = M.make_matrix(np.atleast_1d(np.asarray(a, dtype=np.float64)))
                    if m.shape != (nt, nt):
                        raise ValueError(f"parameter matrix shape {m.shape} != ({nt}, {nt})")
                    return np.ascontiguousarray(m.T)  # column-major storage

                d = _lib.BlockDesc()
                d.kind = _KIND[kind]
                d.ntype = nt
                d.n = n
                d.type_start = t0
                d.horizon = float(b["horizon"])
                ps = float(b["particle_size"]) if b.get("particle_size") else M.particle_size(vol)
                d.particle_size = ps
                d.max_neigh = int(b.get("max_neigh", 100))
                s0 = fmat(b["critical_stretch"])
                arrs = [x, y, v, vol, typ, dens, s0]
                d.x, d.y, d.v, d.volume, d.type = [_ptr(a) for a in (x, y, v, vol, typ)]
                d.density, d.critical_stretch = _ptr(dens), _ptr(s0)
                if kind == "bb":
                    c = np.ascontiguousarray(M.bond_stiffness(M.make_matrix(np.atleast_1d(
                        np.asarray(b["bulk_modulus"], dtype=np.float64))), d.horizon).T)
                    arrs.append(c)
                    d.bond_stiffness = _ptr(c)
                else:
                    K = fmat(b["bulk_modulus"])
                    G = fmat(b["shear_modulus"])
                    arrs += [K, G]
                    d.bulk_modulus, d.shear_modulus = _ptr(K), _ptr(G)
                    if kind == "ep":
                        has_ep = True
                        psi = np.ascontiguousarray(M.ep_psi(M.make_matrix(np.atleast_1d(
                            np.asarray(b["yield_stress"], dtype=np.float64))), d.horizon).T)
                        arrs.append(psi)
                        d.psi = _ptr(psi)
                d.deformed = 0 if np.array_equal(y, x) else 1
                keep.append(arrs)
                ctx.check(L.pd_env_add_block(ctx.h, env_id, ctypes.byref(d), None))
                firsts.append(N)
                ns.append(n)
                N += n
                psizes.append(ps)
                type_ranges.append((t0, t1))
            for cspec in spec.get("contacts", []):
                pair = cspec["pair"]
                # one type list: self contact of one block (ContactModel11,
                # contacts.jl:105-119), else a pair (ContactModel12, :76-91)
                self_contact = len(pair) == 1 or len(pair[1]) == 0
                t1 = np.ascontiguousarray(np.asarray(pair[0], dtype=np.int64).ravel())
                t2 = np.ascontiguousarray(np.asarray([] if self_contact else pair[1], dtype=np.int64).ravel())
                b1 = next(i for i, r in enumerate(type_ranges) if r[0] <= t1[0] <= r[1])
                b2 = b1 if self_contact else next(i for i, r in enumerate(type_ranges) if r[0] <= t2[0] <= r[1])
                dD = float(cspec.get("distanceD", 1.0))
                dX = float(cspec.get("distanceX", 5.0 if self_contact else 3.0))
                equi = psizes[b1] * dD if self_contact else (psizes[b1] + psizes[b2]) / 2.0 * dD
                n2 = ns[b2]
                expo = cspec.get("exponent", 1)
                cd = _lib.ContactDesc()
                cd.types1, cd.ntypes1 = _ptr(t1), t1.size
                cd.types2, cd.ntypes2 = (_ptr(t2), t2.size) if t2.size else (None, 0)
                cd.exponent = float(expo)
                cd.exponent_is_int = int(pow_by_squaring(expo))
                cd.stiffness = float(cspec["stiffness"])
                cd.equi_dist = equi
                cd.distance = dX * equi
                cd.max_neighs = min(int(cspec.get("max_neighs", 50)), n2)
                keep.append([t1, t2])
                ctx.check(L.pd_env_add_contact(ctx.h, env_id, ctypes.byref(cd)))
            for bc in spec.get("bcs", []):
                mask = np.ascontiguousarray(np.asarray(bc["mask"], dtype=np.uint8))
                if mask.size != N:
                    raise ValueError("BC mask must cover every env particle")
                bd = _lib.BcDesc()
                bd.mask = _ptr(mask)
                if bc["kind"] == "fix":
                    bd.kind = _lib.PD_BC_FIX
                else:
                    bd.kind = _lib.PD_BC_TOFRO
                    r = np.asarray(bc["rate"], dtype=np.float64)
                    for i in range(3):
                        bd.rate[i] = float(r[i])
                    bd.freq = float(bc.get("freq", math.inf))
                    bd.applyafter = int(bc.get("applyafter", 0))
                # onlyatstart / dims keywords of the BC constructors (FixBC.jl:46, ToFroBC.jl:51)
                opts = _lib.PD_BC_ONLYATSTART if bc.get("onlyatstart", False) else 0
                dims = bc.get("dims")
                if dims is not None:
                    opts |= _lib.PD_BC_DIMS(sum(1 << k for k in range(3) if dims[k]))
                bd.options = opts
                keep.append([mask])
                ctx.check(L.pd_env_add_bc(ctx.h, env_id, ctypes.byref(bd)))
            ctx.check(L.pd_env_commit(ctx.h, env_id))
        except Exception:
            L.pd_env_destroy(ctx.h, env_id)
            raise
        return cls(ctx, env_id, len(spec["blocks"]), N, firsts, ns, has_ep, spec)

    def close(self):
        if self.id is not None and self.ctx.h:
            self.ctx.lib.pd_env_destroy(self.ctx.h, self.id)
        self.id = None

    # -- stepping -------------------------------------------------------------
    def run(self, nsteps: int, nuf: int = 1):
        ids = np.array([self.id], dtype=np.int64)
        self.ctx.check(self.ctx.lib.pd_run(self.ctx.h, _ptr(ids), 1, int(nsteps), int(nuf)))

    def run_logged(self, nsteps: int, nuf: int = 1, every: int = 1) -> np.ndarray:
        """run() plus the printdata rows (time, px, py, pz, Fx, Fy, Fz, mean
        damage) of step 1 and every `every`-th step: array (rows, 8)."""
        return run_batch_logged([self], nsteps, nuf, every)[:, 0, :]

    def minimize(self, step_size: float, drag: float, max_iter: int = 1000, x_tol: float = 1e-3,
                 f_tol: float = 1e-3):
        """QSDrag minimize! (minimize.jl:105-231).  Returns (iterations, last
        max |f|, last max |dy|)."""
        it, fm, xm = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.ctx.lib.pd_env_minimize(self.ctx.h, self.id, float(step_size), float(drag),
                                                    int(max_iter), float(x_tol), float(f_tol), ctypes.byref(it),
                                                    ctypes.byref(fm), ctypes.byref(xm)))
        return it.value, fm.value, xm.value

    def checkpoint(self) -> bytes:
        """Resident device state as an opaque blob (pd_env_state_save)."""
        from . import snapshot
        return snapshot.checkpoint(self)

    def restore(self, blob: bytes):
        """Loads a checkpoint of an env of the same shape, form and schedule."""
        from . import snapshot
        snapshot.restore(self, blob)

    def observables(self) -> np.ndarray:
        """The 8 printdata columns of the current state (momentum from the
        current velocities, as after update_acc!)."""
        out = np.zeros(8)
        self.ctx.check(self.ctx.lib.pd_env_observables(self.ctx.h, self.id, _ptr(out)))
        return out

    def update_acc(self, y: Optional[np.ndarray] = None) -> np.ndarray:
        f = np.zeros((self.N, 3))
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self.ctx.check(self.ctx.lib.pd_env_update_acc(self.ctx.h, self.id, _ptr(yy), _ptr(f)))
        return f

    def force_density(self, block: int, y: np.ndarray, f: np.ndarray):
        lo = self.block_first[block] + 1
        hi = self.block_first[block] + self.block_n[block]
        self.ctx.check(self.ctx.lib.pd_force_density(self.ctx.h, self.id, block, _ptr(y), _ptr(f), lo, hi))

    def contact_update(self, cidx: int = 0, y: Optional[np.ndarray] = None):
        """update_contact_neighs! of contact model cidx (pd_contact_update)."""
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self.ctx.check(self.ctx.lib.pd_contact_update(self.ctx.h, self.id, int(cidx), _ptr(yy)))

    def contact_neighs(self, cidx: int = 0, max_neighs: int = 0) -> np.ndarray:
        """RM.neighs of contact model cidx as (n1, max_neighs) int64 (= Julia
        max_neighs x n1), 1-based, descending, 0-padded."""
        c = self.spec["contacts"][cidx]
        types1 = c["pair"][0]
        if len(c["pair"]) > 1 and len(c["pair"][1]) > 0:
            types2 = c["pair"][1]
            n1 = int(sum(int(np.isin(np.asarray(b["type"]), types1).sum()) for b in self.spec["blocks"]))
            n2 = int(sum(int(np.isin(np.asarray(b["type"]), types2).sum()) for b in self.spec["blocks"]))
        else:
            n1 = n2 = int(sum(int(np.isin(np.asarray(b["type"]), types1).sum()) for b in self.spec["blocks"]))
        M = max_neighs or max(1, min(int(c.get("max_neighs", 50)), n2))
        out = np.zeros((n1, M), dtype=np.int64)
        self.ctx.check(self.ctx.lib.pd_contact_neighs(self.ctx.h, self.id, int(cidx), _ptr(out), M))
        return out

    def contact_force(self, f: np.ndarray, cidx: int = 0, y: Optional[np.ndarray] = None) -> np.ndarray:
        """short_range_contact! of contact model cidx: f + its terms (pd_contact_force)."""
        out = np.ascontiguousarray(f, dtype=np.float64).copy()
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self.ctx.check(self.ctx.lib.pd_contact_force(self.ctx.h, self.id, int(cidx), _ptr(yy), _ptr(out)))
        return out

    def update_neighs(self):
        self.ctx.check(self.ctx.lib.pd_env_update_neighs(self.ctx.h, self.id))

    # -- state ------------------------------------------------------------------
    def download(self):
        y, v, f = (np.zeros((self.N, 3)) for _ in range(3))
        self.ctx.check(self.ctx.lib.pd_env_download(self.ctx.h, self.id, _ptr(y), _ptr(v), _ptr(f)))
        return y, v, f

    def upload(self, y=None, v=None, f=None):
        a = [None if q is None else np.ascontiguousarray(q, dtype=np.float64) for q in (y, v, f)]
        self.ctx.check(self.ctx.lib.pd_env_upload(self.ctx.h, self.id, *[_ptr(q) for q in a]))

    def family(self, block: int = 0):
        rows = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.pd_block_family_rows(self.ctx.h, self.id, block, ctypes.byref(rows)))
        n = self.block_n[block]
        fam = np.zeros((n, rows.value), dtype=np.int64)
        intact = np.zeros((n, rows.value), dtype=np.uint8)
        self.ctx.check(self.ctx.lib.pd_block_family_export(self.ctx.h, self.id, block, _ptr(fam), _ptr(intact),
                                                           rows.value))
        return fam, intact

    def weighted_volume(self, block: int = 0) -> np.ndarray:
        m = np.zeros(self.block_n[block])
        self.ctx.check(self.ctx.lib.pd_block_weighted_volume(self.ctx.h, self.id, block, _ptr(m)))
        return m

    def theta(self, block: int = 0) -> np.ndarray:
        t = np.zeros(self.block_n[block])
        self.ctx.check(self.ctx.lib.pd_block_theta(self.ctx.h, self.id, block, _ptr(t)))
        return t

    def edp(self, block: int = 0) -> np.ndarray:
        rows = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.pd_block_family_rows(self.ctx.h, self.id, block, ctypes.byref(rows)))
        e = np.zeros((self.block_n[block], rows.value))
        self.ctx.check(self.ctx.lib.pd_block_edp(self.ctx.h, self.id, block, _ptr(e), rows.value))
        return e

    def damage(self) -> np.ndarray:
        d = np.zeros(self.N)
        self.ctx.check(self.ctx.lib.pd_env_damage(self.ctx.h, self.id, _ptr(d)))
        return d

    def info(self):
        n, b, by = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.ctx.lib.pd_env_info(self.ctx.h, self.id, ctypes.byref(n), ctypes.byref(b),
                                                ctypes.byref(by)))
        return {"particles": n.value, "bonds": b.value, "device_bytes": by.value}

    def layout(self) -> dict:
        """Bond layout: lanes per particle, tiled FP64 passes, compact (FP32-exact
        x and volume) halo records, largest tile halo."""
        g, t, h = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(self.ctx.lib.pd_env_layout(self.ctx.h, self.id, ctypes.byref(g), ctypes.byref(t),
                                                  ctypes.byref(h)))
        return {"lanes_per_particle": g.value, "tiled": bool(t.value), "compact_records": t.value == 2,
                "max_halo": h.value}

    @property
    def lanes_per_particle(self) -> int:
        g = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.pd_env_schedule(self.ctx.h, self.id, ctypes.byref(g)))
        return g.value

    @property
    def time_step(self) -> int:
        t = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.pd_env_time_step(self.ctx.h, self.id, ctypes.byref(t)))
        return t.value

    def check_finite(self):
        self.ctx.check(self.ctx.lib.pd_env_check_finite(self.ctx.h, self.id))


def run_batch(envs: Iterable[Env], nsteps: int, nuf: int = 1):
    """Step several envs of one context together (one launch per kernel class)."""
    envs = list(envs)
    ctx = envs[0].ctx
    ids = np.array([e.id for e in envs], dtype=np.int64)
    ctx.check(ctx.lib.pd_run(ctx.h, _ptr(ids), len(envs), int(nsteps), int(nuf)))


def run_batch_begin(envs: Iterable[Env], nsteps: int, nuf: int = 1):
    """Start run_batch and return once its steps are enqueued (pd_run_begin).
    Until run_batch_end, the context accepts only env construction, commit,
    info and the destruction of envs that the run does not step."""
    envs = list(envs)
    ctx = envs[0].ctx
    ids = np.array([e.id for e in envs], dtype=np.int64)
    ctx.check(ctx.lib.pd_run_begin(ctx.h, _ptr(ids), len(envs), int(nsteps), int(nuf)))


def run_batch_done(ctx: Context) -> bool:
    """Whether the steps of the run started by run_batch_begin have completed."""
    done = ctypes.c_int32(0)
    ctx.check(ctx.lib.pd_run_query(ctx.h, ctypes.byref(done)))
    return bool(done.value)


def run_batch_end(ctx: Context):
    """Wait for the run started by run_batch_begin and complete it."""
    ctx.check(ctx.lib.pd_run_end(ctx.h))


def run_batch_logged(envs: Iterable[Env], nsteps: int, nuf: int = 1, every: int = 1) -> np.ndarray:
    """run_batch plus the printdata rows of every env: array (rows, nenv, 8)."""
    envs = list(envs)
    ctx = envs[0].ctx
    ids = np.array([e.id for e in envs], dtype=np.int64)
    rows = int(ctx.lib.pd_log_rows(int(nsteps), int(every)))
    log = np.zeros((max(rows, 1), len(envs), 8))
    ctx.check(ctx.lib.pd_run_logged(ctx.h, _ptr(ids), len(envs), int(nsteps), int(nuf), int(every), _ptr(log)))
    return log[:rows]
