"""Snapshot I/O from device-resident state.

* write_ovito(filename, **columns): the OVITO .data text writer of
  src/io/ff_ovito.jl:81-116 -- first line "N ", second line "# " plus the
  column names sorted and joined by " | ", then one line per particle with
  every column's value(s) in sorted-name order, space separated, each line
  ending in " \\n".  Numbers print as Julia's `string(::Float64)` does:
  shortest round-trip digits, plain notation for decimal exponents in
  [-4, 16), otherwise "d.ddde-5" style.
* save_state(filename, env) / save_state_ovito_bc(filename, env): save_state!
  and save_state_ovito_bc! (src/io/io.jl:44-90), including their update_acc!
  call before writing (which recomputes f and may break bonds); the columns are
  id, type, position, velocity, acceleration, mass, volume, damage.  A ".data"
  name with more than 500,000 particles switches to the key/value dump unless
  force=True, as the reference does (it writes JLD2 there; here the same keys go
  to a NumPy .npz, since no HDF5 library is part of this package).
* checkpoint(env) / restore(env, blob): the resident device state as an opaque
  blob (pd_env_state_save / pd_env_state_load) -- the restart role of
  save_to_file / load_from_file (io.jl:170-187).
"""
from __future__ import annotations

import ctypes
import math
import os
import warnings

import numpy as np


def julia_float(x: float) -> str:
    """`string(x::Float64)` of Julia: shortest round-trip digits (the digits
    Python's repr chooses) in Julia's spelling ("1.0e-5", "2.5e16", "NaN")."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    r = repr(float(x))
    if "e" in r:
        m, e = r.split("e")
        if "." not in m:
            m += ".0"
        return f"{m}e{int(e)}"
    return r


def _fmt(v) -> str:
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return julia_float(float(v))


def write_ovito(filename: str, **columns):
    """ff_ovito.jl:81-116.  Each column is a vector (N,) or a d x N matrix (the
    reference's layout: column i holds particle i); `id` sets N."""
    names = sorted(columns)
    ida = np.asarray(columns["id"])
    N = ida.shape[-1] if ida.ndim > 1 else ida.shape[0]
    cols = [np.asarray(columns[k]) for k in names]
    header = "# " + " | ".join(names)
    with open(filename, "w") as f:
        f.write(f"{N} \n{header}\n")
        for j in range(N):
            vals = []
            for a in cols:
                if a.ndim == 1:
                    vals.append(_fmt(a[j].item() if hasattr(a[j], "item") else a[j]))
                else:
                    vals.extend(_fmt(v.item()) for v in a[:, j])
            f.write(" ".join(vals) + " \n")


def _env_columns(env) -> dict:
    spec = env.spec
    types, mass, vol = [], [], []
    for b in spec["blocks"]:
        t = np.asarray(b["type"], dtype=np.int64)
        v = np.asarray(b["vol"], dtype=np.float64)
        dens = np.atleast_1d(np.asarray(b["density"], dtype=np.float64))
        t0 = int(t.min())
        m = v.copy()
        for k in range(len(dens)):
            mk = t == t0 + k
            m[mk] = m[mk] * dens[k]
        types.append(t)
        mass.append(m)
        vol.append(v)
    y, v, f = env.download()
    return {"type": np.concatenate(types), "position": y.T, "velocity": v.T, "acceleration": f.T,
            "mass": np.concatenate(mass), "volume": np.concatenate(vol)}


def _write_data(filename: str, cols: dict, force: bool):
    N = cols["id"].shape[0]
    if filename.endswith(".data") and N > 500000 and not force:
        filename = filename[:-5] + ".npz"
        warnings.warn("Saving to a key/value dump instead of a data file. Size of the data is too large.")
    if filename.endswith(".data"):
        write_ovito(filename, **cols)
    else:
        np.savez(filename if filename.endswith(".npz") else filename + ".npz", **cols)
    return filename


def save_state(filename: str, env, force: bool = False):
    """save_state! (io.jl:44-54): update_acc!, then id, type, position,
    velocity, acceleration, mass, volume and damage of every particle."""
    env.update_acc()
    cols = _env_columns(env)
    cols["id"] = np.arange(1, env.N + 1, dtype=np.int64)
    cols["damage"] = env.damage()
    return _write_data(filename, cols, force)


def save_state_ovito_bc(filename: str, env, force: bool = False):
    """save_state_ovito_bc! (io.jl:70-90): as save_state with the type column
    replaced by the (1-based) index of the last boundary condition holding
    each particle, 99 for none, into BC_<basename>."""
    env.update_acc()
    cols = _env_columns(env)
    bc_type = np.full(env.N, 99, dtype=np.int64)
    for i, bc in enumerate(env.spec.get("bcs", []), start=1):
        bc_type[np.asarray(bc["mask"], dtype=bool)] = i
    cols["type"] = bc_type
    cols["id"] = np.arange(1, env.N + 1, dtype=np.int64)
    cols["damage"] = env.damage()
    filename = os.path.join(os.path.dirname(filename), "BC_" + os.path.basename(filename))
    return _write_data(filename, cols, force)


SAVE_STATE_DIMS = {"acceleration": 3, "position": 3, "velocity": 3}


def read_ovito(filename: str, dims: dict = SAVE_STATE_DIMS) -> dict:
    """Reads a file written by write_ovito back into columns: (N,) vectors, or
    d x N matrices for the names in `dims` (values per particle)."""
    with open(filename) as f:
        N = int(f.readline().split()[0])
        names = [s.strip() for s in f.readline()[1:].split("|")]
        rows = [line.split() for line in f]
    if len(rows) != N:
        raise ValueError(f"{filename}: {len(rows)} rows for N = {N}")
    out, k = {}, 0
    for n in names:
        d = dims.get(n, 1)
        conv = int if n in ("id", "type") else float
        arr = np.array([[conv(r[k + i]) for r in rows] for i in range(d)])
        out[n] = arr[0] if d == 1 else arr
        k += d
    return out


def checkpoint(env) -> bytes:
    """Resident device state of env as an opaque blob (pd_env_state_save)."""
    ctx = env.ctx
    n = ctypes.c_int64()
    ctx.check(ctx.lib.pd_env_state_size(ctx.h, env.id, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    ctx.check(ctx.lib.pd_env_state_save(ctx.h, env.id, buf, n.value))
    return buf.raw


def restore(env, blob: bytes):
    """Loads a checkpoint into env (same shape, form and schedule)."""
    ctx = env.ctx
    buf = ctypes.create_string_buffer(blob, len(blob))
    ctx.check(ctx.lib.pd_env_state_load(ctx.h, env.id, buf, len(blob)))
