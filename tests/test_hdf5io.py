"""Mesh files in the reference's HDF5 format (readMesh2dHDF5, utilsIO.jl:107-316).

The reader and writer are checked against each other on generated meshes and,
where this container has them, against h5py (run in /opt/conda's python) on
the mesh the reference ships and on files the writer produced."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import pkg

REF_MESH = "/root/reference/oblickShock2dl00F1n.h5"
CONDA_PY = "/opt/conda/bin/python3.9"
FIELDS = ("nCells", "nNodes", "nNeibCells", "nBSets", "xNodes", "yNodes", "mesh_connectivity", "bc_data",
          "bc_indexes", "cell_nodes_X", "cell_nodes_Y", "cell_mid_points", "cell_areas", "HX", "cell_edges_Nx",
          "cell_edges_Ny", "cell_edges_length", "cell_stiffness", "cell_clusters", "node_stencils", "node2cellsL2up",
          "node2cellsL2down", "cells2nodes")


def _same_mesh(a, b):
    for k in FIELDS:
        x, y = getattr(a, k), getattr(b, k)
        if isinstance(x, (int, np.integer)):
            assert int(x) == int(y), k
        else:
            x, y = np.asarray(x), np.asarray(y)
            assert x.shape == y.shape, (k, x.shape, y.shape)
            assert np.array_equal(x, y), k


@pytest.mark.parametrize("quads", ["none", "mixed", "all"])
def test_round_trip_generated_mesh(tmp_path, quads):
    P = pkg()
    m = P.mesh.make_box_mesh(7, 5, 2.0, 1.0, quads=quads)
    fn = str(tmp_path / "box.h5")
    P.write_mesh_hdf5(fn, m)
    r = P.read_mesh_hdf5(fn)
    _same_mesh(m, r)
    assert r.cell_stiffness.dtype == np.int32 and r.node_stencils.dtype == np.float64
    assert r.cell_stiffness.flags.f_contiguous


def test_round_trip_many_datasets_and_dtypes(tmp_path):
    P = pkg()
    rng = np.random.default_rng(3)
    d = {f"a{k:02d}": rng.standard_normal((k % 5 + 1, 7)) for k in range(40)}
    d["i8"] = np.arange(10, dtype=np.int64)
    d["u16"] = np.arange(6, dtype=np.uint16).reshape(2, 3)
    d["f32"] = rng.standard_normal(9).astype(np.float32)
    d["s"] = np.int32(-7)
    fn = str(tmp_path / "many.h5")
    P.hdf5io.write_hdf5(fn, d)
    r = P.hdf5io.read_hdf5(fn)
    assert sorted(r) == sorted(d)
    for k, v in d.items():
        assert np.asarray(v).dtype == r[k].dtype and np.array_equal(np.asarray(v), r[k]), k


def test_rejects_non_hdf5(tmp_path):
    fn = tmp_path / "x.h5"
    fn.write_bytes(b"not hdf5" * 20)
    with pytest.raises(ValueError):
        pkg().hdf5io.read_hdf5(str(fn))


def _h5py(code, *args):
    if not os.path.exists(CONDA_PY):
        pytest.skip("no h5py interpreter in this container")
    r = subprocess.run([CONDA_PY, "-c", code, *args], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("h5py unavailable: " + r.stderr[-200:])
    return r.stdout


def test_reads_the_reference_mesh_like_h5py(tmp_path):
    if not os.path.exists(REF_MESH):
        pytest.skip("reference mesh not available")
    out = tmp_path / "ref.npz"
    _h5py("import h5py, numpy as np, sys\n"
          "f = h5py.File(sys.argv[1], 'r')\n"
          "np.savez(sys.argv[2], **{k: np.array(f[k]) for k in f.keys()})\n", REF_MESH, str(out))
    want = np.load(out)
    got = pkg().hdf5io.read_hdf5(REF_MESH)
    assert sorted(got) == sorted(want.files)
    for k in want.files:
        assert got[k].dtype == want[k].dtype and got[k].shape == want[k].shape and np.array_equal(got[k], want[k]), k
    m = pkg().read_mesh_hdf5(REF_MESH)
    assert (m.nCells, m.nNodes) == (3838, 2147)
    assert m.cell_stiffness.shape == (3838, 4) and m.cell_clusters.shape == (2147, 8)


def test_written_file_reads_in_h5py(tmp_path):
    P = pkg()
    m = P.mesh.make_box_mesh(6, 4, quads="mixed")
    fn = str(tmp_path / "box.h5")
    P.write_mesh_hdf5(fn, m)
    txt = _h5py("import h5py, json, sys\n"
                "f = h5py.File(sys.argv[1], 'r')\n"
                "print(json.dumps({k: [list(f[k].shape), str(f[k].dtype), f[k][()].ravel().tolist()] "
                "for k in f.keys()}))\n", fn)
    d = json.loads(txt)
    assert d["nCells"][2] == [m.nCells] and d["nCells"][1] == "int32"
    st = d["cell_stiffness"]
    assert st[0] == [4, m.nCells] and st[1] == "int32"
    assert np.array_equal(np.array(st[2]).reshape(4, m.nCells).T, m.cell_stiffness)
    ns = d["node_stencils"]
    assert ns[1] == "float64" and np.array_equal(np.array(ns[2]).reshape(8, m.nNodes).T, m.node_stencils)
    assert "node2cellL2up" in d and "node2cellL2down" in d


def test_solution_snapshot_round_trip(tmp_path):
    P = pkg()
    rng = np.random.default_rng(5)
    r, u, v, p = (rng.random(50) + 0.5 for _ in range(4))
    fn = str(tmp_path / "snap.h5")
    P.hdf5io.save_solution_hdf5(fn, 1e-3, 0.25, r, u, v, p)
    d = P.hdf5io.load_solution_hdf5(fn)
    assert d["dt"] == 1e-3 and d["flowTime"] == 0.25 and d["nCells"] == 50
    for k, a in zip(("densityCells", "UxCells", "UyCells", "pressureCells"), (r, u, v, p)):
        assert np.array_equal(d[k], a)


@pytest.mark.gpu
def test_solver_restart_continues(tmp_path):
    """Ten steps, snapshot, ten more, against a solver restarted from the
    snapshot for the last ten (ns FP64, fixed dt).  The snapshot holds
    primitives, as the reference's does, so the restarted conservative state
    agrees to rounding, not bit for bit."""
    from helpers import random_fields
    P = pkg()
    L = P._lib
    m = P.mesh.make_box_mesh(12, 8, quads="mixed")
    f = random_fields(m, rho=(0.8, 1.4), vel=(-0.6, 0.9), p=(0.8, 1.6))
    bcs = {code: (L.GD_BC_EXTRAPOLATE, ()) for code in (-1, -2, -3, -4)}
    prm = P.make_params(scheme=L.GD_SCHEME_NS, viscous=True, RGAS=1.0, mu_ref=2e-3, T_ref=1.0, S_suth=0.4, bcs=bcs)
    a = P.GodunovSolver(m, prm)
    a.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells)
    a.run(10, dt=1e-3)
    fn = str(tmp_path / "snap.h5")
    a.save_snapshot(fn, dt=1e-3)
    a.run(10, dt=1e-3)
    b = P.GodunovSolver(m, prm)
    b.load_snapshot(fn)
    assert abs(b.flow_time - (a.flow_time - 1e-2)) < 1e-15
    b.run(10, dt=1e-3)
    x, y = a.conservative(), b.conservative()
    assert np.allclose(x, y, rtol=1e-12, atol=1e-12 * np.abs(x).max())
    a.close()
    b.close()
