"""Node fields as VTK XML unstructured grids (saveResults*VTK, utilsIO.jl:2-53)."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import pkg


@pytest.mark.parametrize("binary", [True, False])
def test_vtu_round_trip_mixed_mesh(tmp_path, binary):
    P = pkg()
    m = P.mesh.make_box_mesh(5, 4, 2.0, 1.0, quads="mixed")
    rng = np.random.default_rng(1)
    f = {k: rng.standard_normal(m.nNodes) for k in ("density", "Ux", "Uy", "pressure")}
    fn = P.vtkio.save_results_vtk(str(tmp_path / "out"), m, f, binary=binary)
    assert fn.endswith(".vtu")
    root = ET.parse(fn).getroot()
    assert root.get("type") == "UnstructuredGrid"
    d = P.vtkio.read_vtu(fn)
    assert np.array_equal(d["points"][:, 0], m.xNodes) and np.array_equal(d["points"][:, 1], m.yNodes)
    assert np.all(d["points"][:, 2] == 0)
    nv = m.mesh_connectivity[:, 2]
    assert np.array_equal(d["offsets"], np.cumsum(nv))
    assert set(np.unique(d["types"])) == {5, 9}
    assert np.array_equal(d["types"] == 9, nv == 4)
    # cell i's nodes, 0-based, in the mesh's vertex order
    for i in (0, m.nCells // 2, m.nCells - 1):
        a, b = (d["offsets"][i - 1] if i else 0), d["offsets"][i]
        assert np.array_equal(d["connectivity"][a:b], m.mesh_connectivity[i, 3:3 + nv[i]] - 1)
    for k, v in f.items():
        assert np.array_equal(d["point_data"][k], v), k


def test_reference_named_writers(tmp_path):
    P = pkg()
    m = P.mesh.make_tgv_periodic_mesh(8)
    r = np.linspace(1, 2, m.nNodes)
    fn = P.saveResults4VTK(str(tmp_path / "s4"), m, r, 0 * r, 1 + r, 2 * r)
    assert sorted(P.vtkio.read_vtu(fn)["point_data"]) == ["Ux", "Uy", "density", "pressure"]
    fn = P.saveResults2VTK(str(tmp_path / "s2"), m, r)
    assert list(P.vtkio.read_vtu(fn)["point_data"]) == ["density"]
    with pytest.raises(ValueError):
        P.vtkio.save_results_vtk(str(tmp_path / "bad"), m, {"x": r[:-1]})


@pytest.mark.gpu
def test_solver_writes_node_primitives(tmp_path):
    """The device state's node interpolation written as VTK equals the
    oracle's node interpolation of the same cell state."""
    import oracle as O
    from helpers import random_fields
    P = pkg()
    L = P._lib
    m = P.mesh.make_box_mesh(6, 5, quads="mixed")
    f = random_fields(m)
    s = P.GodunovSolver(m, P.make_params(), order=L.GD_ORDER_RCM)
    s.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells)
    fn = s.save_vtk(str(tmp_path / "state"))
    s.close()
    d = P.vtkio.read_vtu(fn)["point_data"]
    U = O.phs2dcns(f, 1.4, m.nCells)
    O.cells2nodes(O.MeshView(m), f, U)
    for name, ref in (("density", f.densityNodes), ("Ux", f.UxNodes), ("Uy", f.UyNodes), ("pressure", f.pressureNodes)):
        assert np.allclose(d[name], ref, rtol=1e-13, atol=1e-13), name
