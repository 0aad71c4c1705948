"""CPU-side checks of the C ABI: the library loads and exports every entry point
include/peridyn.h declares, and the descriptor structs match the header layout."""
import ctypes
import os
import re

from peridyn_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "peridyn.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pd_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(L, name), name


def test_struct_sizes_match_header_layout():
    # 4+4+8+8 + 5 pointers + 3 doubles/ints + 6 pointers + 4+4
    assert ctypes.sizeof(_lib.BlockDesc) == 8 * 2 + 8 + 5 * 8 + 8 + 8 + 8 + 6 * 8 + 8
    assert ctypes.sizeof(_lib.ContactDesc) == 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8
    assert ctypes.sizeof(_lib.BcDesc) == 8 + 8 + 24 + 8 + 8


def test_julia_binding_calls_only_exported_symbols_and_defines_its_seams():
    """julia/PeriDynHIP.jl cannot be run here (no Julia), so its text is checked:
    every entry point it ccalls is declared in include/peridyn.h, and the seam
    methods its docstring lists are defined."""
    src = open(os.path.join(ROOT, "julia", "PeriDynHIP.jl")).read()
    called = set(re.findall(r"ccall\(\(:(pd_\w+), LIB\)", src))
    assert called and called <= set(header_functions()), called - set(header_functions())
    for seam in ("function PeriDyn.force_density_T!(", "function PeriDyn.short_range_contact!(",
                 "function PeriDyn.update_contact_neighs!(", "function register!(", "function copy_intact!("):
        assert seam in src, seam
    assert "pd_contact_force" in called and "pd_contact_update" in called and "pd_contact_neighs" in called
