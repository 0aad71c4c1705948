"""ctypes binding of libperidyn_hip.so (the C ABI in include/peridyn.h).

The product path has no CPU fallback: if the HIP library is missing or no HIP
device is present, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libperidyn_hip.so")

PD_OK, PD_E_HIP, PD_E_ARG, PD_E_OVERFLOW, PD_E_STATE, PD_E_NOMEM, PD_E_NONFINITE = 0, -1, -2, -3, -4, -5, -6
PD_BB, PD_OSB, PD_EP = 0, 1, 2
PD_PREC_EXACT, PD_PREC_FP64, PD_PREC_FP32C = 0, 1, 2
PD_BC_FIX, PD_BC_TOFRO = 0, 1
PD_BC_ONLYATSTART = 1


def PD_BC_DIMS(d: int) -> int:
    return (int(d) & 7) << 8

EXPORTS = [
    "pd_ctx_create", "pd_ctx_destroy", "pd_last_error", "pd_sync",
    "pd_env_create", "pd_env_add_block", "pd_env_add_contact", "pd_env_add_bc", "pd_env_commit",
    "pd_env_destroy", "pd_env_info", "pd_block_family_rows", "pd_block_family_export",
    "pd_block_weighted_volume", "pd_block_theta", "pd_block_edp", "pd_run", "pd_env_update_acc",
    "pd_force_density", "pd_env_update_neighs", "pd_env_download", "pd_env_upload", "pd_env_damage",
    "pd_env_time_step", "pd_env_check_finite", "pd_ctx_profile", "pd_ctx_profile_read",
    "pd_env_set_schedule", "pd_env_schedule", "pd_env_layout", "pd_ctx_mem_info", "pd_run_logged", "pd_log_rows",
    "pd_env_observables", "pd_env_state_size", "pd_env_state_save", "pd_env_state_load",
    "pd_env_minimize", "pd_env_observables_noalloc", "pd_contact_update", "pd_contact_neighs",
    "pd_contact_force", "pd_run_begin", "pd_run_query", "pd_run_end",
]
PROF_CLASSES = {"prologue": 0, "latch": 1, "theta": 2, "td": 3, "force_bb": 4, "force_osb": 5,
                "force_ep": 6, "contact": 7, "book": 8}

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double


class BlockDesc(ctypes.Structure):
    _fields_ = [("kind", I32), ("ntype", I32), ("n", I64), ("type_start", I64),
                ("x", P), ("y", P), ("v", P), ("volume", P), ("type", P),
                ("horizon", D), ("particle_size", D), ("max_neigh", I64),
                ("density", P), ("critical_stretch", P), ("bond_stiffness", P),
                ("bulk_modulus", P), ("shear_modulus", P), ("psi", P),
                ("deformed", I32), ("reserved", I32)]


class ContactDesc(ctypes.Structure):
    _fields_ = [("types1", P), ("ntypes1", I64), ("types2", P), ("ntypes2", I64),
                ("exponent", D), ("exponent_is_int", I32), ("reserved", I32),
                ("stiffness", D), ("equi_dist", D), ("distance", D), ("max_neighs", I64)]


class BcDesc(ctypes.Structure):
    _fields_ = [("kind", I32), ("options", I32), ("mask", P), ("rate", D * 3),
                ("freq", D), ("applyafter", I64)]


class PeriDynError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load the in-tree HIP library (fails loudly when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m peridyn_amd.build` "
                          "(no CPU fallback exists for the hot path)")
    L = ctypes.CDLL(LIB_PATH)
    for name in EXPORTS:
        getattr(L, name)  # raises AttributeError if a symbol is missing
    L.pd_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
    L.pd_ctx_destroy.argtypes = [P]
    L.pd_last_error.argtypes = [P]
    L.pd_last_error.restype = ctypes.c_char_p
    L.pd_sync.argtypes = [P]
    L.pd_ctx_mem_info.argtypes = [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]
    L.pd_run_logged.argtypes = [P, P, I64, I64, I64, I64, P]
    L.pd_log_rows.argtypes = [I64, I64]
    L.pd_log_rows.restype = I64
    L.pd_env_observables.argtypes = [P, I64, P]
    L.pd_env_observables_noalloc.argtypes = [P, I64, P]
    L.pd_env_state_size.argtypes = [P, I64, ctypes.POINTER(I64)]
    L.pd_env_state_save.argtypes = [P, I64, P, I64]
    L.pd_env_state_load.argtypes = [P, I64, P, I64]
    L.pd_env_minimize.argtypes = [P, I64, D, D, I64, D, D, ctypes.POINTER(I64), ctypes.POINTER(D),
                                  ctypes.POINTER(D)]
    L.pd_env_create.argtypes = [P, D, I32, ctypes.POINTER(I64)]
    L.pd_env_add_block.argtypes = [P, I64, ctypes.POINTER(BlockDesc), ctypes.POINTER(I64)]
    L.pd_env_add_contact.argtypes = [P, I64, ctypes.POINTER(ContactDesc)]
    L.pd_env_add_bc.argtypes = [P, I64, ctypes.POINTER(BcDesc)]
    L.pd_env_commit.argtypes = [P, I64]
    L.pd_env_destroy.argtypes = [P, I64]
    L.pd_env_info.argtypes = [P, I64, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]
    L.pd_block_family_rows.argtypes = [P, I64, I64, ctypes.POINTER(I64)]
    L.pd_block_family_export.argtypes = [P, I64, I64, P, P, I64]
    L.pd_block_weighted_volume.argtypes = [P, I64, I64, P]
    L.pd_block_theta.argtypes = [P, I64, I64, P]
    L.pd_block_edp.argtypes = [P, I64, I64, P, I64]
    L.pd_run.argtypes = [P, P, I64, I64, I64]
    L.pd_run_begin.argtypes = [P, P, I64, I64, I64]
    L.pd_run_query.argtypes = [P, ctypes.POINTER(I32)]
    L.pd_run_end.argtypes = [P]
    L.pd_env_update_acc.argtypes = [P, I64, P, P]
    L.pd_force_density.argtypes = [P, I64, I64, P, P, I64, I64]
    L.pd_env_update_neighs.argtypes = [P, I64]
    L.pd_contact_update.argtypes = [P, I64, I64, P]
    L.pd_contact_neighs.argtypes = [P, I64, I64, P, I64]
    L.pd_contact_force.argtypes = [P, I64, I64, P, P]
    L.pd_env_download.argtypes = [P, I64, P, P, P]
    L.pd_env_upload.argtypes = [P, I64, P, P, P]
    L.pd_env_damage.argtypes = [P, I64, P]
    L.pd_env_time_step.argtypes = [P, I64, ctypes.POINTER(I64)]
    L.pd_env_check_finite.argtypes = [P, I64]
    L.pd_env_set_schedule.argtypes = [P, I64, I32]
    L.pd_env_schedule.argtypes = [P, I64, ctypes.POINTER(I32)]
    L.pd_env_layout.argtypes = [P, I64, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32)]
    L.pd_ctx_profile.argtypes = [P, I32]
    L.pd_ctx_profile_read.argtypes = [P, I32, ctypes.POINTER(D), ctypes.POINTER(I64)]
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc != PD_OK:
        msg = load().pd_last_error(ctx)
        raise PeriDynError(rc, (msg or b"").decode(errors="replace"))
