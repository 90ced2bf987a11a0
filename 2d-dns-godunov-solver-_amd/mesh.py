"""Mesh data model and generators in the reference's ``mesh2d_Int32`` format.

The reference keeps its unstructured 2-D mesh in one immutable struct
(``primeObjects.jl:109-135``) filled by ``readMesh2dHDF5`` (``utilsIO.jl:107-316``).
Every ``(nCells, k)`` array is Julia column-major, i.e. one contiguous vector per
column; here they are numpy arrays in Fortran order with the same shapes, the
same 1-based ids and the same conventions:

* ``mesh_connectivity[i,:] = [id, type (3 tri / 2 quad), nVertices, v1..v4]``
* edge ``k`` of cell ``i`` is ``cells2nodes[i, 2k-1:2k]``; tri edges are
  ``(v1v2, v2v3, v1v3)``, quad edges ``(v1v2, v2v3, v3v4, v1v4)``
* ``cell_edges_Nx/Ny`` are unit **inward** normals, ``cell_edges_length`` the
  edge length (column 4 is zero for triangles)
* ``cell_stiffness[i,k]`` is the neighbour id across edge ``k`` (>0) or a
  boundary code (<0)
* ``node2cellsL2down[i,2k-1]`` is the own cell's vertex opposite edge ``k``;
  ``node2cellsL2up[i,2k-1]`` is the neighbour's vertex opposite the shared edge
  (quads use the pair ``2k-1, 2k`` of far-side nodes)
* ``cell_clusters[J,:]`` lists the cells touching node ``J`` (zero padded to
  ``nNeibCells`` = 8) and ``node_stencils[J,:]`` their interpolation weights
  (1.0 in padded slots)

The reference ships no mesh generator, and its periodic TGV and cylinder meshes
are absent, so the generators below are this package's own.  The node weights
are inverse centroid distances (the reference's weight formula is not in its
sources; both the oracle and the device path read the same array).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Optional

import numpy as np

N_NEIB_CELLS = 8


@dataclasses.dataclass
class THERMOPHYSICS:
    """Gas constants, ``thermo.jl:4-10`` / ``initfields_TGV2d.jl:217-223``."""
    RGAS: float = 287.058
    Gamma: float = 1.4
    Cp: float = 1000.0
    mu: float = 0.0
    kGas: float = 0.0


@dataclasses.dataclass
class mesh2d_Int32:  # noqa: N801  (reference name)
    """Mirror of ``mesh2d_Int32`` (``primeObjects.jl:109-135``)."""
    nCells: int
    nNodes: int
    nNeibCells: int
    nBSets: int
    xNodes: np.ndarray
    yNodes: np.ndarray
    mesh_connectivity: np.ndarray
    bc_data: np.ndarray
    bc_indexes: np.ndarray
    cell_nodes_X: np.ndarray
    cell_nodes_Y: np.ndarray
    cell_mid_points: np.ndarray
    cell_areas: np.ndarray
    HX: np.ndarray
    cell_edges_Nx: np.ndarray
    cell_edges_Ny: np.ndarray
    cell_edges_length: np.ndarray
    cell_stiffness: np.ndarray
    cell_clusters: np.ndarray
    node_stencils: np.ndarray
    node2cellsL2up: np.ndarray
    node2cellsL2down: np.ndarray
    cells2nodes: np.ndarray
    periodic: bool = False

    def n_boundary_faces(self) -> int:
        return int(np.count_nonzero(self.cell_stiffness[:, :3] < 0)) + int(
            np.count_nonzero((self.cell_stiffness[:, 3] < 0) & (self.mesh_connectivity[:, 2] == 4)))


@dataclasses.dataclass
class fields2d:  # noqa: N801  (reference name)
    """Mirror of ``fields2d`` (``primeObjects.jl:85-106``)."""
    densityCells: np.ndarray
    UxCells: np.ndarray
    UyCells: np.ndarray
    pressureCells: np.ndarray
    aSoundCells: np.ndarray
    VMAXCells: np.ndarray
    densityNodes: np.ndarray
    UxNodes: np.ndarray
    UyNodes: np.ndarray
    pressureNodes: np.ndarray

    @staticmethod
    def zeros(n_cells: int, n_nodes: int) -> "fields2d":
        z = lambda n: np.zeros(n, dtype=np.float64)  # noqa: E731
        return fields2d(z(n_cells), z(n_cells), z(n_cells), z(n_cells), z(n_cells), z(n_cells),
                        z(n_nodes), z(n_nodes), z(n_nodes), z(n_nodes))

    def copy(self) -> "fields2d":
        return fields2d(*[np.array(getattr(self, f.name), copy=True) for f in dataclasses.fields(self)])


# ----------------------------------------------------------------------------
# generic builder
# ----------------------------------------------------------------------------

def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def build_mesh(x_nodes: np.ndarray, y_nodes: np.ndarray, cell_vertices: np.ndarray,
               cell_x: np.ndarray, cell_y: np.ndarray,
               boundary_code: Optional[Callable[[np.ndarray, np.ndarray, np.ndarray, np.ndarray], np.ndarray]] = None,
               periodic: bool = False) -> mesh2d_Int32:
    """Build the reference mesh arrays from cells.

    ``cell_vertices`` is ``(Nc, 4)`` 0-based node ids with ``-1`` in the 4th
    column for triangles; ``cell_x/cell_y`` are the vertex coordinates as seen
    from the cell (for periodic meshes they are unwrapped, i.e. they may differ
    from ``x_nodes`` by a period).  ``boundary_code(xm, ym, nx_out, ny_out)``
    returns the (negative) code of each boundary edge from its midpoint and
    outward normal.
    """
    cv = np.array(cell_vertices, dtype=np.int64)
    cx = np.array(cell_x, dtype=np.float64)
    cy = np.array(cell_y, dtype=np.float64)
    nc = cv.shape[0]
    nn = x_nodes.shape[0]
    nv = np.where(cv[:, 3] >= 0, 4, 3)

    # orient counter-clockwise (shoelace on the cell's own coordinates)
    def signed_area(cv_, cx_, cy_, nv_):
        a = np.zeros(nc)
        for k in range(4):
            k1 = np.where(nv_ == 4, (k + 1) % 4, (k + 1) % 3)
            valid = k < nv_
            xk = cx_[np.arange(nc), k]
            yk = cy_[np.arange(nc), k]
            xk1 = cx_[np.arange(nc), k1]
            yk1 = cy_[np.arange(nc), k1]
            a += np.where(valid, xk * yk1 - xk1 * yk, 0.0)
        return 0.5 * a

    sa = signed_area(cv, cx, cy, nv)
    flip = sa < 0
    if np.any(flip):
        # reverse v2..vn keeps v1 first
        for arr in (cv, cx, cy):
            tri = flip & (nv == 3)
            quad = flip & (nv == 4)
            arr[tri, 1], arr[tri, 2] = arr[tri, 2].copy(), arr[tri, 1].copy()
            arr[quad, 1], arr[quad, 3] = arr[quad, 3].copy(), arr[quad, 1].copy()
        sa = signed_area(cv, cx, cy, nv)
    if np.any(sa <= 0):
        raise ValueError("degenerate cell")
    area = sa

    # edge k -> local vertex pair (a, b); tri (0,1),(1,2),(0,2); quad (0,1),(1,2),(2,3),(0,3)
    ea = np.array([[0, 1, 0, -1], [0, 1, 2, 0]])  # row 0 tri, row 1 quad
    eb = np.array([[1, 2, 2, -1], [1, 2, 3, 3]])
    # far-side local vertices: tri opposite vertex; quad the two nodes not on the edge
    far1 = np.array([[2, 0, 1, -1], [2, 3, 0, 1]])
    far2 = np.array([[2, 0, 1, -1], [3, 0, 1, 2]])
    isq = (nv == 4).astype(np.int64)
    ar = np.arange(nc)

    mid_x = np.where(nv == 4, cx.sum(axis=1) / 4.0, cx[:, :3].sum(axis=1) / 3.0)
    mid_y = np.where(nv == 4, cy.sum(axis=1) / 4.0, cy[:, :3].sum(axis=1) / 3.0)

    nx = np.zeros((nc, 4)); ny = np.zeros((nc, 4)); ln = np.zeros((nc, 4))
    stiff = np.zeros((nc, 4), dtype=np.int64)
    c2n = np.zeros((nc, 8), dtype=np.int64)
    down = np.zeros((nc, 8), dtype=np.int64)
    up = np.zeros((nc, 8), dtype=np.int64)

    # half edges
    he_cell, he_k, he_key = [], [], []
    for k in range(4):
        valid = k < nv
        la = ea[isq, k]; lb = eb[isq, k]
        la = np.where(valid, la, 0); lb = np.where(valid, lb, 0)
        a_id = cv[ar, la]; b_id = cv[ar, lb]
        xa, ya = cx[ar, la], cy[ar, la]
        xb, yb = cx[ar, lb], cy[ar, lb]
        tx, ty = xb - xa, yb - ya
        L = np.hypot(tx, ty)
        n1x, n1y = -ty / np.where(L > 0, L, 1.0), tx / np.where(L > 0, L, 1.0)
        xm, ym = 0.5 * (xa + xb), 0.5 * (ya + yb)
        s = np.sign(n1x * (mid_x - xm) + n1y * (mid_y - ym))
        s = np.where(s == 0, 1.0, s)
        nx[:, k] = np.where(valid, n1x * s, 0.0)
        ny[:, k] = np.where(valid, n1y * s, 0.0)
        ln[:, k] = np.where(valid, L, 0.0)
        c2n[:, 2 * k] = np.where(valid, a_id + 1, 0)
        c2n[:, 2 * k + 1] = np.where(valid, b_id + 1, 0)
        f1 = cv[ar, np.where(valid, far1[isq, k], 0)]
        f2 = cv[ar, np.where(valid, far2[isq, k], 0)]
        down[:, 2 * k] = np.where(valid, f1 + 1, 0)
        down[:, 2 * k + 1] = np.where(valid & (nv == 4), f2 + 1, 0)
        idx = np.nonzero(valid)[0]
        he_cell.append(idx)
        he_k.append(np.full(idx.shape, k))
        he_key.append(np.minimum(a_id, b_id)[idx] * nn + np.maximum(a_id, b_id)[idx])
    he_cell = np.concatenate(he_cell); he_k = np.concatenate(he_k); he_key = np.concatenate(he_key)
    order = np.argsort(he_key, kind="stable")
    ks = he_key[order]
    same_next = np.zeros(ks.shape, dtype=bool)
    same_next[:-1] = ks[:-1] == ks[1:]
    if np.any(same_next[:-1] & same_next[1:]):
        raise ValueError("non-manifold edge (shared by more than two cells)")
    first = np.nonzero(same_next)[0]
    second = first + 1
    paired = np.zeros(ks.shape, dtype=bool)
    paired[first] = True; paired[second] = True
    c1, k1 = he_cell[order[first]], he_k[order[first]]
    c2, k2 = he_cell[order[second]], he_k[order[second]]
    stiff[c1, k1] = c2 + 1
    stiff[c2, k2] = c1 + 1
    # the neighbour's far-side node(s) across the shared edge
    for (ca, ka, cb, kb) in ((c1, k1, c2, k2), (c2, k2, c1, k1)):
        up[ca, 2 * ka] = down[cb, 2 * kb]
        second_far = down[cb, 2 * kb + 1]
        up[ca, 2 * ka + 1] = np.where(nv[ca] == 4, np.where(second_far > 0, second_far, down[cb, 2 * kb]), 0)
    bidx = order[~paired]
    if bidx.size:
        bc, bk = he_cell[bidx], he_k[bidx]
        if boundary_code is None:
            raise ValueError("mesh has boundary edges but no boundary_code given")
        la = ea[isq[bc], bk]; lb = eb[isq[bc], bk]
        xm = 0.5 * (cx[bc, la] + cx[bc, lb]); ym = 0.5 * (cy[bc, la] + cy[bc, lb])
        codes = np.asarray(boundary_code(xm, ym, -nx[bc, bk], -ny[bc, bk]), dtype=np.int64)
        if np.any(codes >= 0):
            raise ValueError("boundary codes must be negative")
        stiff[bc, bk] = codes

    # node clusters: cells touching each node, ascending cell id
    inc_node = []; inc_cell = []; inc_slot = []
    for k in range(4):
        valid = k < nv
        idx = np.nonzero(valid)[0]
        inc_node.append(cv[idx, k]); inc_cell.append(idx); inc_slot.append(np.full(idx.shape, k))
    inc_node = np.concatenate(inc_node); inc_cell = np.concatenate(inc_cell); inc_slot = np.concatenate(inc_slot)
    o = np.lexsort((inc_cell, inc_node))
    inc_node, inc_cell, inc_slot = inc_node[o], inc_cell[o], inc_slot[o]
    counts = np.bincount(inc_node, minlength=nn)
    if counts.max() > N_NEIB_CELLS:
        raise ValueError("a node touches more than 8 cells")
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pos = np.arange(inc_node.size) - start[inc_node]
    clusters = np.zeros((nn, N_NEIB_CELLS), dtype=np.int64)
    stencils = np.ones((nn, N_NEIB_CELLS), dtype=np.float64)
    clusters[inc_node, pos] = inc_cell + 1
    dx = cx[inc_cell, inc_slot] - mid_x[inc_cell]
    dy = cy[inc_cell, inc_slot] - mid_y[inc_cell]
    stencils[inc_node, pos] = 1.0 / np.hypot(dx, dy)

    conn = np.zeros((nc, 7), dtype=np.int64)
    conn[:, 0] = np.arange(1, nc + 1)
    conn[:, 1] = np.where(nv == 4, 2, 3)
    conn[:, 2] = nv
    conn[:, 3:7] = np.where(cv >= 0, cv + 1, 0)
    cnx = np.where(cv >= 0, cx, 0.0); cny = np.where(cv >= 0, cy, 0.0)

    i32 = np.int32
    return mesh2d_Int32(
        nCells=nc, nNodes=nn, nNeibCells=N_NEIB_CELLS, nBSets=int(len(np.unique(stiff[stiff < 0]))),
        xNodes=np.asarray(x_nodes, dtype=np.float64).copy(), yNodes=np.asarray(y_nodes, dtype=np.float64).copy(),
        mesh_connectivity=_fortran(conn, i32),
        bc_data=np.zeros((0, 3), dtype=i32, order="F"), bc_indexes=np.zeros(0, dtype=i32),
        cell_nodes_X=_fortran(cnx, np.float64), cell_nodes_Y=_fortran(cny, np.float64),
        cell_mid_points=_fortran(np.stack([mid_x, mid_y], axis=1), np.float64),
        cell_areas=np.ascontiguousarray(area), HX=np.ascontiguousarray(ln.max(axis=1)),
        cell_edges_Nx=_fortran(nx, np.float64), cell_edges_Ny=_fortran(ny, np.float64),
        cell_edges_length=_fortran(ln, np.float64), cell_stiffness=_fortran(stiff, i32),
        cell_clusters=_fortran(clusters, i32), node_stencils=_fortran(stencils, np.float64),
        node2cellsL2up=_fortran(up, i32), node2cellsL2down=_fortran(down, i32),
        cells2nodes=_fortran(c2n, i32), periodic=periodic)


# ----------------------------------------------------------------------------
# generators
# ----------------------------------------------------------------------------

def make_tgv_periodic_mesh(n: int = 100, length: float = 2.0 * math.pi) -> mesh2d_Int32:
    """Doubly periodic ``[0, length]^2`` square, ``n x n`` quads each cut along
    the SW-NE diagonal into (SW, SE, NE) and (SW, NE, NW): ``2 n^2`` triangles
    and ``n^2`` unique nodes; every face is interior."""
    h = length / n
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xn = (ii * h).ravel(); yn = (jj * h).ravel()
    nid = lambda i, j: (j % n) * n + (i % n)  # noqa: E731
    i = ii.ravel(); j = jj.ravel()
    sw, se, ne, nw = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
    x0, x1 = i * h, (i + 1) * h
    y0, y1 = j * h, (j + 1) * h
    nq = n * n
    cv = np.full((2 * nq, 4), -1, dtype=np.int64)
    cx = np.zeros((2 * nq, 4)); cy = np.zeros((2 * nq, 4))
    cv[0::2, :3] = np.stack([sw, se, ne], 1)
    cx[0::2, :3] = np.stack([x0, x1, x1], 1); cy[0::2, :3] = np.stack([y0, y0, y1], 1)
    cv[1::2, :3] = np.stack([sw, ne, nw], 1)
    cx[1::2, :3] = np.stack([x0, x1, x0], 1); cy[1::2, :3] = np.stack([y0, y1, y1], 1)
    return build_mesh(xn, yn, cv, cx, cy, None, periodic=True)


def make_box_mesh(nx_: int, ny_: int, lx: float = 1.0, ly: float = 1.0, quads: str = "none",
                  codes=(-1, -2, -3, -4), x0: float = 0.0, y0: float = 0.0) -> mesh2d_Int32:
    """Non-periodic rectangle split SW-NE into triangles.

    ``quads``: ``"none"`` (all triangles), ``"all"`` (all quads) or ``"mixed"``
    (left half quads, right half triangles).  Boundary codes are
    ``codes = (top, right, bottom, left)`` like the shipped oblique-shock mesh.
    """
    hx, hy = lx / nx_, ly / ny_
    jj, ii = np.meshgrid(np.arange(ny_ + 1), np.arange(nx_ + 1), indexing="ij")
    xn = (x0 + ii * hx).ravel(); yn = (y0 + jj * hy).ravel()
    nid = lambda i, j: j * (nx_ + 1) + i  # noqa: E731
    cvs, cxs, cys = [], [], []
    for j in range(ny_):
        for i in range(nx_):
            sw, se, ne, nw = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            X = [x0 + i * hx, x0 + (i + 1) * hx]; Y = [y0 + j * hy, y0 + (j + 1) * hy]
            as_quad = quads == "all" or (quads == "mixed" and i < nx_ // 2)
            if as_quad:
                cvs.append([sw, se, ne, nw]); cxs.append([X[0], X[1], X[1], X[0]]); cys.append([Y[0], Y[0], Y[1], Y[1]])
            else:
                cvs.append([sw, se, ne, -1]); cxs.append([X[0], X[1], X[1], 0]); cys.append([Y[0], Y[0], Y[1], 0])
                cvs.append([sw, ne, nw, -1]); cxs.append([X[0], X[1], X[0], 0]); cys.append([Y[0], Y[1], Y[1], 0])
    top, right, bottom, left = codes
    eps = 1e-9 * max(lx, ly)

    def bcode(xm, ym, nxo, nyo):
        out = np.full(xm.shape, left, dtype=np.int64)
        out[nyo > 0.5] = top
        out[nxo > 0.5] = right
        out[nyo < -0.5] = bottom
        return out
    del eps
    return build_mesh(xn, yn, np.array(cvs), np.array(cxs), np.array(cys), bcode)


def ogrid_radii(n_r: int, r0: float, r1: float, dr0: float) -> np.ndarray:
    """Radial node positions ``r0 .. r1`` with geometric stretching: first
    spacing ``dr0``, ratio ``q`` solved from ``dr0 (q^(n-1) - 1)/(q - 1) = r1 - r0``."""
    m = n_r - 1
    span = r1 - r0
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        q = 0.5 * (lo + hi)
        lqm = m * np.log(q)  # q^m overflows for fine grids at the upper bracket
        s = np.inf if lqm > 700.0 else dr0 * np.expm1(lqm) / (q - 1.0)
        if s > span:
            hi = q
        else:
            lo = q
    q = 0.5 * (lo + hi)
    r = r0 + dr0 * (q ** np.arange(n_r) - 1.0) / (q - 1.0)
    r[-1] = r1
    return r


CODE_INFLOW = -1
CODE_OUTFLOW = -2
CODE_WALL = -3


def make_cylinder_ogrid_mesh(n_theta: int = 720, n_r: int = 1445, r0: float = 0.5, r1: float = 10.0,
                             dr0: float = 2e-4, flow_dir=(1.0, 0.0)) -> mesh2d_Int32:
    """O-grid around a cylinder of radius ``r0``: ``n_theta`` uniform
    circumferential nodes (periodic in theta) x ``n_r`` geometrically stretched
    radial nodes; each quad is split into two triangles, giving
    ``2 n_theta (n_r - 1)`` cells.  Boundary codes: cylinder wall -3; outer
    circle -1 (inflow) where the free stream enters, -2 (outflow) elsewhere."""
    r = ogrid_radii(n_r, r0, r1, dr0)
    th = 2.0 * math.pi * np.arange(n_theta) / n_theta
    R, TH = np.meshgrid(r, th, indexing="ij")  # (n_r, n_theta)
    xn = (R * np.cos(TH)).ravel(); yn = (R * np.sin(TH)).ravel()
    j, i = np.meshgrid(np.arange(n_r - 1), np.arange(n_theta), indexing="ij")
    j = j.ravel(); i = i.ravel()
    ip = (i + 1) % n_theta
    a = j * n_theta + i; b = j * n_theta + ip; c = (j + 1) * n_theta + ip; d = (j + 1) * n_theta + i
    nq = a.size
    cv = np.full((2 * nq, 4), -1, dtype=np.int64)
    cv[0::2, :3] = np.stack([a, c, b], 1)
    cv[1::2, :3] = np.stack([a, d, c], 1)
    cx = np.zeros((2 * nq, 4)); cy = np.zeros((2 * nq, 4))
    cx[:, :3] = xn[np.maximum(cv[:, :3], 0)]; cy[:, :3] = yn[np.maximum(cv[:, :3], 0)]
    rmid = 0.5 * (r0 + r1)
    fx, fy = flow_dir

    def bcode(xm, ym, nxo, nyo):
        rr = np.hypot(xm, ym)
        out = np.where(rr < rmid, CODE_WALL, np.where(nxo * fx + nyo * fy < 0.0, CODE_INFLOW, CODE_OUTFLOW))
        return out.astype(np.int64)
    return build_mesh(xn, yn, cv, cx, cy, bcode)


# ----------------------------------------------------------------------------
# renumbering
# ----------------------------------------------------------------------------

_MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """``n`` outputs of the splitmix64 generator (Steele, Lea, Flood 2014)
    started at ``seed``; a portable, documented RNG for fixtures."""
    with np.errstate(over="ignore"):
        s = (np.uint64(seed & _MASK64) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64))
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Doubles in (0, 1) from the top 53 bits of splitmix64."""
    return ((splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """Standard normal samples by Box-Muller on splitmix64 uniforms."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]; out[1::2] = z[m:]
    return out[:n]


def random_permutation(n: int, seed: int) -> np.ndarray:
    """Seeded permutation ``new -> old`` (argsort of splitmix64 keys)."""
    return np.argsort(splitmix64(seed, n), kind="stable")


def cell_adjacency(mesh: mesh2d_Int32):
    """CSR adjacency of cells through interior faces (0-based)."""
    st = np.asarray(mesh.cell_stiffness, dtype=np.int64)
    nv = np.asarray(mesh.mesh_connectivity[:, 2])
    rows, cols = [], []
    for k in range(4):
        ok = (st[:, k] > 0) & (k < nv)
        idx = np.nonzero(ok)[0]
        rows.append(idx); cols.append(st[idx, k] - 1)
    rows = np.concatenate(rows); cols = np.concatenate(cols)
    return rows, cols


def rcm_permutation(mesh: mesh2d_Int32) -> np.ndarray:
    """Reverse Cuthill-McKee order of the cell graph, ``new -> old``."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    rows, cols = cell_adjacency(mesh)
    a = sp.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(mesh.nCells, mesh.nCells))
    return np.asarray(reverse_cuthill_mckee(a, symmetric_mode=True), dtype=np.int64)


def permute_cells(mesh: mesh2d_Int32, perm: np.ndarray) -> mesh2d_Int32:
    """Renumber cells: new cell ``c`` is old cell ``perm[c]`` (0-based).  Node
    numbering is unchanged; every cell id stored in the mesh is rewritten."""
    perm = np.asarray(perm, dtype=np.int64)
    nc = mesh.nCells
    inv = np.empty(nc, dtype=np.int64)
    inv[perm] = np.arange(nc)

    def rows(a):
        return np.asfortranarray(np.asarray(a)[perm])
    st = np.array(mesh.cell_stiffness, dtype=np.int64)[perm]
    st = np.where(st > 0, inv[np.maximum(st, 1) - 1] + 1, st)
    cl = np.array(mesh.cell_clusters, dtype=np.int64)
    cl = np.where(cl > 0, inv[np.maximum(cl, 1) - 1] + 1, cl)
    conn = np.array(mesh.mesh_connectivity)[perm]
    conn[:, 0] = np.arange(1, nc + 1)
    i32 = np.int32
    return dataclasses.replace(
        mesh,
        mesh_connectivity=_fortran(conn, i32), cell_nodes_X=rows(mesh.cell_nodes_X), cell_nodes_Y=rows(mesh.cell_nodes_Y),
        cell_mid_points=rows(mesh.cell_mid_points), cell_areas=np.ascontiguousarray(mesh.cell_areas[perm]),
        HX=np.ascontiguousarray(mesh.HX[perm]), cell_edges_Nx=rows(mesh.cell_edges_Nx),
        cell_edges_Ny=rows(mesh.cell_edges_Ny), cell_edges_length=rows(mesh.cell_edges_length),
        cell_stiffness=_fortran(st, i32), cell_clusters=_fortran(cl, i32),
        node2cellsL2up=rows(mesh.node2cellsL2up), node2cellsL2down=rows(mesh.node2cellsL2down),
        cells2nodes=rows(mesh.cells2nodes))


def distibuteCellsInThreadsSA(nThreads: int, nCells: int) -> np.ndarray:  # noqa: N802 (reference name)
    """Static contiguous partition, ``partMesh2d.jl:2-34`` (1-based, inclusive).

    ``floor(nCells/nThreads)`` cells per thread, the remainder goes to thread 1;
    with one thread the reference returns the all-zero table (kept)."""
    ct = np.zeros((nThreads, 2), dtype=np.int32)
    if nThreads > 1:
        n_parts = nCells // nThreads
        for i in range(1, nThreads + 1):
            ct[i - 1, 1] = nCells - n_parts * (nThreads - i)
        for i in range(1, nThreads + 1):
            ct[i - 1, 0] = ct[i - 1, 1] - n_parts + 1
        ct[0, 0] = 1
    return ct


distibuteNodesInThreadsSA = distibuteCellsInThreadsSA
