"""The C ABI's structs as the binding declares them: every ctypes struct of
treeamr_hip._abi has the size and the field offsets that a C compiler gives
the same struct in include/tah.h (a tiny C program, compiled with gcc, prints
them). A field added on one side only, or in another order, fails here
before any call passes a struct across the boundary."""
import os
import shutil
import subprocess

import pytest

from treeamr_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = ["tah_operators", "tah_stencil1d", "tah_group_desc", "tah_schedule_desc", "tah_layout",
           "tah_plan_options", "tah_boundary_region", "tah_family_options"]


def _c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tah.h"', 'int main(void) {']
    for name in STRUCTS:
        cls = getattr(_abi, name)
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, key, value = line.split()
        got[(name, key)] = int(value)
    return got


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to compile the header")
def test_ctypes_structs_match_the_header(tmp_path):
    import ctypes as C
    c = _c_layout(tmp_path)
    for name in STRUCTS:
        cls = getattr(_abi, name)
        assert C.sizeof(cls) == c[(name, "size")], name
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == c[(name, field)], (name, field)
