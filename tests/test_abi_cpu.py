"""CPU-side checks of the C ABI: the library loads and exports every symbol the header declares."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "radiant_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int64_t|int|const char \*)\s*\*?\s*(rt_\w+)\s*\(", txt, re.M)))


def test_header_matches_binding():
    from radiant_hip import _lib
    assert set(header_symbols()) == set(_lib.EXPORTS)


def test_library_exports_all_symbols():
    from radiant_hip import _lib
    L = _lib.lib()
    for s in header_symbols():
        assert hasattr(L, s), s


def test_struct_sizes_match_c():
    """ctypes layouts against the C layout (sizes computed from the declared field order)."""
    from radiant_hip import _lib
    assert C.sizeof(_lib.GroupStats) == 8 + 8 + 8 + 4 * 4
    assert C.sizeof(_lib.ServerConfig) == 16
