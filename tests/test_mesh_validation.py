"""gd_mesh_create validates a mesh on the host before any device call, so its
error paths are checked here without a GPU: every corrupted mesh is refused
with GD_ERR_INVALID and a message naming the problem, and a valid one gets as
far as the device (GD_ERR_NO_DEVICE on a machine without one)."""
import ctypes as C

import numpy as np
import pytest

from conftest import pkg


def _create(mesh, order=0, perm=None):
    L = pkg()._lib
    lib = L.load()
    keep = []

    def d(a):
        a = np.asfortranarray(a, dtype=np.float64); keep.append(a); return L.ptr(a)

    def i(a):
        a = np.asfortranarray(a, dtype=np.int32); keep.append(a); return L.ptr(a, C.c_int32)
    desc = L.gd_mesh_desc(mesh.nCells, mesh.nNodes, int(mesh.cell_clusters.shape[1]), 0, d(mesh.xNodes),
                          d(mesh.yNodes), i(mesh.mesh_connectivity), d(mesh.cell_nodes_X), d(mesh.cell_nodes_Y),
                          d(mesh.cell_mid_points), d(mesh.cell_areas), d(mesh.HX), d(mesh.cell_edges_Nx),
                          d(mesh.cell_edges_Ny), d(mesh.cell_edges_length), i(mesh.cell_stiffness),
                          i(mesh.cell_clusters), d(mesh.node_stencils), i(mesh.node2cellsL2up),
                          i(mesh.node2cellsL2down), i(mesh.cells2nodes))
    h = C.c_void_p()
    pp = None
    if perm is not None:
        perm = np.ascontiguousarray(perm, dtype=np.int64)
        keep.append(perm)
        pp = perm.ctypes.data_as(C.POINTER(C.c_int64))
    rc = lib.gd_mesh_create(C.byref(desc), order, pp, C.byref(h))
    if rc == 0:
        lib.gd_mesh_destroy(h)
    return rc, L.last_error()


def _box():
    return pkg().mesh.make_box_mesh(6, 5, quads="mixed")


@pytest.mark.parametrize("order", ["GD_ORDER_AS_GIVEN", "GD_ORDER_RCM", "GD_ORDER_HILBERT"])
def test_valid_mesh_reaches_the_device(order):
    L = pkg()._lib
    rc, msg = _create(_box(), order=getattr(L, order))
    assert rc in (0, L.GD_ERR_NO_DEVICE), msg


def test_unknown_order_is_refused():
    L = pkg()._lib
    rc, msg = _create(_box(), order=7)
    assert rc == L.GD_ERR_INVALID and "unknown cell order" in msg


@pytest.mark.parametrize("case, expect", [
    ("vertices", "neither 3 nor 4"),
    ("neighbour_range", "neighbour id out of range"),
    ("asymmetric", "not symmetric"),
    ("area", "non-positive cell area"),
    ("edge_nodes", "cells2nodes out of range"),
    ("cluster", "cell_clusters id out of range"),
    ("duplicate_neighbour", "not symmetric"),  # the displaced neighbour no longer finds the cell
])
def test_corrupted_mesh_is_refused(case, expect):
    L = pkg()._lib
    m = _box()
    if case == "vertices":
        m.mesh_connectivity[3, 2] = 5
    elif case == "neighbour_range":
        m.cell_stiffness[0, 0] = m.nCells + 7
    elif case == "asymmetric":
        i = int(np.argmax((m.cell_stiffness[:, :3] > 0).all(axis=1)))
        j = m.cell_stiffness[i, 0] - 1
        k = list(m.cell_stiffness[j]).index(i + 1)
        m.cell_stiffness[j, k] = -2  # j no longer sees i
    elif case == "area":
        m.cell_areas[2] = 0.0
    elif case == "edge_nodes":
        m.cells2nodes[1, 0] = m.nNodes + 1
    elif case == "cluster":
        m.cell_clusters[0, 0] = m.nCells + 3
    elif case == "duplicate_neighbour":
        # cell i lists neighbour j on two faces (and j lists i twice)
        i = int(np.argmax((m.cell_stiffness[:, :3] > 0).all(axis=1) & (m.mesh_connectivity[:, 2] == 3)))
        j = m.cell_stiffness[i, 0]
        m.cell_stiffness[i, 1] = j
    rc, msg = _create(m)
    assert rc == L.GD_ERR_INVALID, (rc, msg)
    assert expect in msg, msg


def test_bad_arguments():
    L = pkg()._lib
    lib = L.load()
    m = _box()
    rc, msg = _create(m, order=L.GD_ORDER_PERMUTATION)
    assert rc == L.GD_ERR_INVALID and "without a permutation" in msg
    rc, msg = _create(m, order=L.GD_ORDER_PERMUTATION, perm=np.zeros(m.nCells, dtype=np.int64))
    assert rc == L.GD_ERR_INVALID and "not a permutation" in msg
    assert lib.gd_mesh_create(None, 0, None, None) == L.GD_ERR_INVALID
    assert lib.gd_run(None, 1, None, None) == L.GD_ERR_INVALID
    assert lib.gd_adjoint_run(None, None, None, None, None, None) == L.GD_ERR_INVALID
    assert lib.gd_state_create(None, None, None) == L.GD_ERR_INVALID
