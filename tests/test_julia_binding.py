"""The committed Julia binding (julia/LibGodunov.jl, LibGodunovChainRules.jl)
cannot run here (no Julia), so its text is checked against the C ABI: every
`ccall` names an exported entry point with as many argument types as the
header declares, every mirrored struct lists the header struct's fields in
order, and every GD_* constant has the header's value."""
import os
import re

from conftest import ROOT, pkg

JL = os.path.join(ROOT, "2d-dns-godunov-solver-_amd", "julia")


def _src(name):
    return open(os.path.join(JL, name)).read()


def _split_top(s):
    """split a ccall argument-type tuple at its top-level commas"""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip()); cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _ccalls(src):
    """(symbol, [argument types]) of each ccall((:sym, libgodunov), Ret, (types...), ...)"""
    out = []
    for m in re.finditer(r"ccall\(\(:(gd_\w+), libgodunov\),\s*\w+,\s*\(", src):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        out.append((m.group(1), _split_top(src[m.end():i - 1])))
    return out


def test_every_ccall_matches_the_c_abi():
    L = pkg()._lib
    lib = L.load()
    calls = _ccalls(_src("LibGodunov.jl"))
    assert len(calls) >= 20
    names = set()
    for name, types in calls:
        assert name in L.EXPORTS, name
        assert len(types) == len(getattr(lib, name).argtypes), (name, types)
        names.add(name)
    for must in ("gd_mesh_create", "gd_state_create", "gd_calc_one_stage", "gd_calc_one_stage_cuda",
                 "gd_calc_face_states", "gd_calc_diff_term", "gd_update_variables", "gd_cells2nodes", "gd_run",
                 "gd_adjoint_run"):
        assert must in names, must


def _julia_fields(src, struct):
    m = re.search(r"(?:mutable )?struct " + struct + r"\n(.*?)\nend", src, re.S)
    assert m, struct
    body = "\n".join(line for line in m.group(1).splitlines() if "=" not in line)
    return re.findall(r"(\w+)::", body)


def test_mirrored_structs_list_the_header_fields_in_order():
    L = pkg()._lib
    src = _src("LibGodunov.jl")
    for jl, ct in (("GdMeshDesc", L.gd_mesh_desc), ("GdBC", L.gd_bc), ("GdParams", L.gd_params),
                   ("GdRunParams", L.gd_run_params), ("GdRunStats", L.gd_run_stats),
                   ("GdAdjointParams", L.gd_adjoint_params), ("GdAdjointStats", L.gd_adjoint_stats)):
        assert _julia_fields(src, jl) == [f[0] for f in ct._fields_], jl


def test_constants_have_the_header_values():
    L = pkg()._lib
    src = _src("LibGodunov.jl")
    n = 0
    for m in re.finditer(r"const ((?:GD_\w+,\s*)*GD_\w+)\s*=\s*(.+)", src):
        names = [x.strip() for x in m.group(1).split(",")]
        rhs = m.group(2).strip()
        if rhs.startswith("Int32.(") and ":" in rhs:
            a, b = (int(x) for x in re.findall(r"(-?\d+):(-?\d+)", rhs)[0])
            vals = list(range(a, b + 1))
        else:
            vals = [int(x) for x in re.findall(r"Int32\((-?\d+)\)", rhs)] or [int(rhs)]
        assert len(vals) == len(names), m.group(0)
        for k, v in zip(names, vals):
            assert getattr(L, k) == v, k
            n += 1
    assert n >= 30


def test_chain_rules_file_uses_the_binding():
    src = _src("LibGodunovChainRules.jl")
    assert "ChainRulesCore.rrule(::typeof(drag_sum)" in src
    assert "drag_gradient" in src and "using ..LibGodunov" in src


def test_snapshot_io_reads_the_datasets_this_package_writes(tmp_path):
    """LibGodunovIO.jl reads and writes exactly the datasets of
    hdf5io.save_solution_hdf5 (the solutionCellsT fields)."""
    g = pkg()
    import numpy as np
    path = str(tmp_path / "snap.h5")
    g.hdf5io.save_solution_hdf5(path, 1e-3, 0.25, np.ones(3), np.zeros(3), np.zeros(3), np.ones(3))
    written = set(g.hdf5io.read_hdf5(path))
    src = _src("LibGodunovIO.jl")
    read = set(re.findall(r'read\(io, "(\w+)"\)', src))
    wrote = set(re.findall(r'write\(io, "(\w+)"', src))
    assert read == written - {"dt"} and wrote == written
