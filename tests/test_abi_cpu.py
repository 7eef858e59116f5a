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


def _c_layout(struct, fields, tmp_path):
    """offsetof() of every field and sizeof() of `struct`, compiled with gcc against the header."""
    import subprocess
    src = tmp_path / "layout.c"
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "radiant_hip.h"', "int main(void) {"]
    lines += [f'  printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += [f'  printf("%zu\\n", sizeof({struct}));', "  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    return vals[:-1], vals[-1]


def test_struct_offsets_match_header(tmp_path):
    """Every field of every ABI struct sits at the C offset in the ctypes binding."""
    from radiant_hip import _lib
    for struct, cls in (("rt_problem_desc", _lib.ProblemDesc), ("rt_group_stats", _lib.GroupStats),
                        ("rt_solve_info", _lib.SolveInfo), ("rt_server_config", _lib.ServerConfig)):
        names = [f[0] for f in cls._fields_]
        offs, size = _c_layout(struct, names, tmp_path)
        got = [getattr(cls, n).offset for n in names]
        assert got == offs, (struct, [(n, a, b) for n, a, b in zip(names, got, offs) if a != b])
        assert C.sizeof(cls) == size, struct


def test_header_fields_all_bound(tmp_path):
    """The ctypes ProblemDesc names every member the header declares in rt_problem_desc."""
    from radiant_hip import _lib
    txt = open(os.path.join(ROOT, "include", "radiant_hip.h")).read()
    body = txt[txt.index("typedef struct rt_problem_desc {"):txt.index("} rt_problem_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    decl = re.findall(r"^\s*(?:const\s+)?\w+\s+([^;]+);", body.split("{", 1)[1], re.M)
    names = set()
    for d in decl:
        for part in d.split(","):
            names.add(re.sub(r"[\s*]|\[\d+\]", "", part))
    assert names == {f[0] for f in _lib.ProblemDesc._fields_}
