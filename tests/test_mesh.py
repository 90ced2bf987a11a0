"""Mesh format conventions (mesh2d_Int32) of the generators, checked on the
generated meshes and, where the container has it, on the mesh the reference
ships (oblickShock2dl00F1n.h5, read with the package's HDF5 reader)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import pkg

REF_MESH = "/root/reference/oblickShock2dl00F1n.h5"


def M():
    return pkg().mesh


def check_conventions(m, periodic=False):
    nc = m.nCells
    nv = m.mesh_connectivity[:, 2]
    assert set(np.unique(nv)) <= {3, 4}
    # unit inward normals: closure sum_k L_k n_k = 0 and n points to the centroid
    for k in range(4):
        ok = k < nv
        L = m.cell_edges_length[ok, k]
        assert np.allclose(np.hypot(m.cell_edges_Nx[ok, k], m.cell_edges_Ny[ok, k]), 1.0, atol=1e-12)
        assert np.all(L > 0)
    sx = (m.cell_edges_Nx * m.cell_edges_length).sum(1)
    sy = (m.cell_edges_Ny * m.cell_edges_length).sum(1)
    assert np.max(np.abs(sx)) < 1e-12 * m.cell_edges_length.max() * 10
    assert np.max(np.abs(sy)) < 1e-12 * m.cell_edges_length.max() * 10
    X, Y = m.cell_nodes_X, m.cell_nodes_Y
    for i in range(0, nc, max(1, nc // 500)):
        n = nv[i]
        # counter-clockwise vertices, area by the shoelace formula
        a = 0.5 * sum(X[i, k] * Y[i, (k + 1) % n] - X[i, (k + 1) % n] * Y[i, k] for k in range(n))
        assert a > 0 and a == pytest.approx(m.cell_areas[i], rel=1e-9)
        for k in range(n):
            e1, e2 = m.cells2nodes[i, 2 * k], m.cells2nodes[i, 2 * k + 1]
            ids = list(m.mesh_connectivity[i, 3:3 + n])
            xa, ya = X[i, ids.index(e1)], Y[i, ids.index(e1)]
            xb, yb = X[i, ids.index(e2)], Y[i, ids.index(e2)]
            cx, cy = m.cell_mid_points[i]
            nx, ny = m.cell_edges_Nx[i, k], m.cell_edges_Ny[i, k]
            assert nx * (cx - 0.5 * (xa + xb)) + ny * (cy - 0.5 * (ya + yb)) > 0      # inward
            assert abs(nx * (xb - xa) + ny * (yb - ya)) < 1e-9 * m.cell_edges_length[i, k]
            j = m.cell_stiffness[i, k]
            if j > 0 and n == 3 and nv[j - 1] == 3:
                # down = own vertex opposite the edge; up = the neighbour's opposite vertex
                d = m.node2cellsL2down[i, 2 * k]
                assert d in ids and d not in (e1, e2)
                u = m.node2cellsL2up[i, 2 * k]
                jd = list(m.mesh_connectivity[j - 1, 3:6])
                assert u in jd and u not in (e1, e2)
    # neighbour symmetry
    st = m.cell_stiffness
    for i in range(0, nc, max(1, nc // 500)):
        for k in range(nv[i]):
            j = st[i, k]
            if j > 0:
                assert (st[j - 1, :nv[j - 1]] == i + 1).sum() == 1


def test_tgv_mesh_counts_and_conventions():
    m = M().make_tgv_periodic_mesh(100)
    assert (m.nCells, m.nNodes) == (20000, 10000)
    assert m.n_boundary_faces() == 0
    check_conventions(m, periodic=True)
    assert np.allclose(m.cell_areas, (2 * np.pi / 100) ** 2 / 2)


def test_cylinder_mesh_counts_and_conventions():
    m = M().make_cylinder_ogrid_mesh(72, 31, 0.5, 10.0, 2e-3)
    assert m.nCells == 2 * 72 * 30 and m.nNodes == 72 * 31
    assert m.n_boundary_faces() == 2 * 72
    st = m.cell_stiffness[:, :3]
    assert (st == -3).sum() == 72 and ((st == -1) | (st == -2)).sum() == 72
    check_conventions(m)
    r = M().ogrid_radii(1445, 0.5, 10.0, 2e-4)
    assert r[1] - r[0] == pytest.approx(2e-4, rel=1e-9) and r[-1] == 10.0
    # full-size count of the benchmark mesh: 720 x 1445 nodes, two triangles per quad
    assert 2 * 720 * 1444 == 2079360


def test_box_mesh_mixed_conventions():
    check_conventions(M().make_box_mesh(9, 7, 2.0, 1.0, quads="mixed"))
    check_conventions(M().make_box_mesh(5, 4, 1.0, 1.0, quads="all"))


def test_permutations():
    m = M().make_box_mesh(10, 8, quads="mixed")
    p = M().random_permutation(m.nCells, 20261203)
    assert sorted(p.tolist()) == list(range(m.nCells))
    mp = M().permute_cells(m, p)
    check_conventions(mp)
    assert np.allclose(mp.cell_areas, m.cell_areas[p])
    r = M().rcm_permutation(mp)
    mr = M().permute_cells(mp, r)
    rows, cols = M().cell_adjacency(mr)
    rows0, cols0 = M().cell_adjacency(mp)
    assert np.abs(rows - cols).max() < np.abs(rows0 - cols0).max()
    # the seeded RNG is stable
    assert M().splitmix64(20261203, 2).tolist() == M().splitmix64(20261203, 2).tolist()
    z = M().normal(20261203, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_step_is_numbering_independent_in_the_oracle():
    """Renumbering cells permutes the result and changes nothing else."""
    from helpers import random_fields
    m = M().make_box_mesh(9, 7, quads="mixed")
    th = M().THERMOPHYSICS(RGAS=1.0)
    f = random_fields(m)
    p = M().random_permutation(m.nCells, 7)
    mp = M().permute_cells(m, p)
    fp = M().fields2d.zeros(m.nCells, m.nNodes)
    for name in ("densityCells", "UxCells", "UyCells", "pressureCells"):
        getattr(fp, name)[:] = getattr(f, name)[p]
    res = []
    for mesh, ff in ((m, f), (mp, fp)):
        mv = O.MeshView(mesh)
        U = O.phs2dcns(ff, 1.4, mesh.nCells)
        O.cells2nodes(mv, ff, U)
        bcs = {-1: (O.BC_FIXED, (1.0, 0.5, 0.0, 1.0)), -3: (O.BC_SLIP_WALL, ())}
        Ud, _ = O.calc_diff_term(mv, ff, th, O.params(bcs=bcs), U)
        res.append(O.calc_one_stage_muscl_cpu_fast(mv, ff, th, O.params(bcs=bcs), U, Ud, 1.0, 1e-3, 0.0)[1])
    assert np.array_equal(res[0][p], res[1])


def _load_reference_mesh(tmp_path):
    if not os.path.exists(REF_MESH):
        pytest.skip("reference mesh not available")
    return pkg().read_mesh_hdf5(REF_MESH)


def test_reference_shipped_mesh_conventions_and_oracle_step(tmp_path):
    """The reference's own mixed tri/quad mesh follows the conventions the
    generators and the oracle assume, and the oracle step preserves a uniform
    state on it (inflow -1 fixed to the same state, -2 zero-gradient, -3 slip
    wall, -4 fixed)."""
    m = _load_reference_mesh(tmp_path)
    assert (m.nCells, m.nNodes) == (3838, 2147)
    assert sorted(np.unique(m.mesh_connectivity[:, 2]).tolist()) == [3, 4]
    check_conventions(m)
    th = M().THERMOPHYSICS(RGAS=287.058)
    f = M().fields2d.zeros(m.nCells, m.nNodes)
    q = (1.0, 290.0, 0.0, 7142.857142857143)
    f.densityCells[:], f.UxCells[:], f.UyCells[:], f.pressureCells[:] = q
    mv = O.MeshView(m)
    U = O.phs2dcns(f, 1.4, m.nCells)
    O.cells2nodes(mv, f, U)
    bcs = {-1: (O.BC_FIXED, q), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_SLIP_WALL, ()), -4: (O.BC_FIXED, q)}
    _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, O.params(bcs=bcs), U, None, 1.0, 1e-7, 0.0)
    wall = (m.cell_stiffness == -3).any(1)
    # away from the wall (where the slip ghost differs from the uniform flow) nothing changes
    assert np.max(np.abs(UN[~wall] - U[~wall]) / np.maximum(np.abs(U).max(0), 1.0)) < 1e-13
