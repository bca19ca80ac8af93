"""The C ABI library: loads, exports every declared entry point, struct
layouts agree with C, program validation (host code, no device needed)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from isosurface_amd import _abi as A
from isosurface_amd.sources import Difference, Sphere, Translate, Union, union_all

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "isosurface.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:const\s+)?\w+\s*\*?\s*(iso_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "iso_extract_mc" in names and "iso_ctx_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = A.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    # and the ctypes table covers them all
    assert {n for n, _, _ in A.SIGNATURES} == set(declared_functions())


STRUCTS = [("iso_node", A.IsoNode), ("iso_grid_desc", A.IsoGridDesc),
           ("iso_mesh_out", A.IsoMeshOut), ("iso_kernel_stat", A.IsoKernelStat),
           ("iso_chunk_desc", A.IsoChunkDesc), ("iso_batch_out", A.IsoBatchOut),
           ("iso_band_train_params", A.IsoBandTrainParams), ("iso_band_stats", A.IsoBandStats)]


def test_struct_layouts_match_c(tmp_path):
    # sizeof and every field's offset of each ctypes mirror against the C header
    lines, want = [], []
    for cname, cls in STRUCTS:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(cls))
        for field in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {field[0]}));')
            want.append(getattr(cls, field[0]).offset)
    src = tmp_path / "layout.c"
    src.write_text('#include "isosurface.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){' + "".join(lines) + 'return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", f"-I{ROOT / 'include'}",
                    str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                       check=True).stdout.split()))
    assert got == want


def test_status_strings():
    lib = A.load()
    for st in range(9):
        assert lib.iso_status_string(st)
    assert lib.iso_abi_version() == 1


def validate(program):
    from isosurface_amd.sources import to_nodes
    lib = A.load()
    msg = C.create_string_buffer(256)
    nodes = to_nodes(program) if program else (A.IsoNode * 1)()
    st = lib.iso_program_validate(nodes, len(program), msg, 256)
    return st, msg.value.decode()


def test_program_validation_accepts_good_programs():
    src = Translate((0.5, 0.5, 0.5), Difference(Sphere(0.1), union_all(
        Translate((0.1 * i, 0.0, 0.0), Sphere(0.05)) for i in range(100))))
    assert validate(src.program()) == (A.ISO_OK, "")


@pytest.mark.parametrize("program,fragment", [
    ([], "empty"),
    ([(A.ISO_OP_UNION, ())], "two operands"),
    ([(A.ISO_OP_SPHERE, (1.0,)), (A.ISO_OP_SPHERE, (1.0,))], "leaves 2 values"),
    ([(A.ISO_OP_TRANSLATE, (0.0, 0.0, 0.0)), (A.ISO_OP_SPHERE, (1.0,))], "unterminated"),
    ([(A.ISO_OP_END_TRANSLATE, ())], "without TRANSLATE"),
    ([(99, ())], "unknown opcode"),
    ([(A.ISO_OP_SPHERE, (1.0,)), (A.ISO_OP_SPHERE, (1.0,)), (A.ISO_OP_SMOOTH_UNION, (0.0,))],
     "k > 0"),
])
def test_program_validation_rejects(program, fragment):
    st, msg = validate(program)
    assert st == A.ISO_ERR_PROGRAM
    assert fragment in msg


def test_program_validation_keeps_right_deep_chains_shallow():
    # Operands are evaluated deeper-first, so right-deep chains need few slots.
    deep = Sphere(1.0)
    for _ in range(40):
        deep = Difference(Union(Sphere(1.0), Sphere(2.0)), deep)
    assert validate(deep.program())[0] == A.ISO_OK
    right = Sphere(1.0)
    for _ in range(40):
        right = Difference(right, Union(Sphere(1.0), Difference(Sphere(0.5), Sphere(2.0))))
    assert validate(right.program())[0] == A.ISO_OK


def balanced(depth):
    if depth == 0:
        return Translate((0.0, 0.0, 0.0), Union(Sphere(1.0), Sphere(0.5)))
    return Difference(balanced(depth - 1), balanced(depth - 1))


def test_program_validation_rejects_too_deep_stacks():
    # A balanced tree of compound operands needs one slot per level.
    assert validate(balanced(7).program())[0] == A.ISO_OK  # 8 slots
    st, msg = validate(balanced(8).program())  # 9 slots
    assert st == A.ISO_ERR_PROGRAM and "value stack" in msg


def test_program_validation_bounds_translation_nesting():
    s = Sphere(1.0)
    for _ in range(6):
        s = Translate((0.1, 0.0, 0.0), Union(s, Sphere(0.5)))
    st, msg = validate(s.program())
    assert st == A.ISO_ERR_PROGRAM and "translations" in msg
