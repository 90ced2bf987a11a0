"""The C-ABI library loads and exports every entry point include/godunov.h declares.

No compute calls here: this runs without a GPU.
"""
import ctypes
import os
import re

from conftest import ROOT, pkg


def declared_functions():
    src = open(os.path.join(ROOT, "include", "godunov.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_binding_list():
    L = pkg()._lib
    assert sorted(L.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    L = pkg()._lib
    lib = L.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_version_and_errors_without_device():
    L = pkg()._lib
    lib = L.load()
    assert b"gfx950" in lib.gd_version()
    buf = ctypes.create_string_buffer(16)
    assert lib.gd_last_error(buf, 16) == 0
    assert lib.gd_last_error(None, 0) == L.GD_ERR_INVALID


def test_struct_sizes_match_header_layout():
    L = pkg()._lib
    # 4 ints, 6 doubles, 2 ints, 8 doubles, 17 x (2 ints + 4 doubles)
    assert ctypes.sizeof(L.gd_params) == 4 * 4 + 6 * 8 + 2 * 4 + 8 * 8 + 17 * (8 + 32)
    assert ctypes.sizeof(L.gd_run_params) == 4 + 4 + 4 * 8 + 4 + 4
    assert ctypes.sizeof(L.gd_mesh_desc) == 8 + 8 + 4 + 4 + 17 * 8
    assert ctypes.sizeof(L.gd_adjoint_params) == 4 + 4 + 8 + 8 + 4 + 4 + 2 * 8 + 8
    assert ctypes.sizeof(L.gd_adjoint_stats) == 8 + 2 * 8 + 8 + 4 + 4 + 8 + 8 + 8


def _c_layout(structs):
    """sizeof and offsetof of every field of the given ctypes structures, as
    the C compiler lays out the header's structs (gcc on a generated file)."""
    import subprocess
    import tempfile
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "godunov.h"', "int main(void) {"]
    for st in structs:
        name = st.__name__
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f in st._fields_:
            lines.append(f'  printf("{name} {f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {tuple(line.split()[:2]): int(line.split()[2]) for line in out.splitlines()}


def test_ctypes_layouts_match_the_c_compiler():
    """Every ctypes mirror of a header struct has the C compiler's size and
    field offsets (a reordering of equal-size fields fails here)."""
    L = pkg()._lib
    structs = [L.gd_mesh_desc, L.gd_mesh_info, L.gd_bc, L.gd_params, L.gd_run_params, L.gd_run_stats,
               L.gd_adjoint_params, L.gd_adjoint_stats]
    c = _c_layout(structs)
    for st in structs:
        assert ctypes.sizeof(st) == c[(st.__name__, "sizeof")], st.__name__
        for f in st._fields_:
            assert getattr(st, f[0]).offset == c[(st.__name__, f[0])], (st.__name__, f[0])


def test_stream_rates_reject_bad_sizes_before_any_device_call():
    L = pkg()._lib
    lib = L.load()
    cp, rd = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    for nbytes, iters in [(1000, 1), (16 * 1024 + 8, 1), (1 << 20, 0)]:
        assert lib.gd_hbm_stream_rates(nbytes, iters, ctypes.byref(cp), ctypes.byref(rd)) == L.GD_ERR_INVALID
    assert lib.gd_hbm_stream_rates(1 << 20, 1, None, ctypes.byref(rd)) == L.GD_ERR_INVALID
    assert cp.value == -1.0 and rd.value == -1.0
