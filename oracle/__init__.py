"""ctypes front end of the CPU oracle: TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the bench's CPU-baseline leg
use this module, and only as the checker.  It wraps ``oracle/liboracle.so``
(built from ``oracle.cpp`` / ``oracle_ns.cpp`` by ``oracle/Makefile``), a scalar
restatement of the reference's Julia step.  Inputs are the reference structs
(``mesh2d_Int32``, ``fields2d``, ``THERMOPHYSICS``) as numpy arrays in the
reference's column-major, 1-based layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_NONE, BC_FIXED, BC_EXTRAPOLATE, BC_SLIP_WALL, BC_TGV_MIRROR, BC_ISOTHERMAL_WALL = range(6)
FLUX_AUSM_PLUS, FLUX_ROE, FLUX_AUSM_PLUS_M2020 = range(3)
LIM_REF_HARMONIC, LIM_VAN_LEER_A, LIM_VAN_ALBADA_A, LIM_MINMOD, LIM_VENKATAKRISHNAN = range(5)
ENERGY_AIR_POLY, ENERGY_CONST = range(2)
MAX_BC = 16

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


class OrcMesh(C.Structure):
    _fields_ = [("nCells", C.c_int64), ("nNodes", C.c_int64), ("nNeibCells", C.c_int32),
                ("xNodes", dp), ("yNodes", dp), ("mesh_connectivity", ip),
                ("cell_nodes_X", dp), ("cell_nodes_Y", dp), ("cell_mid_points", dp),
                ("cell_areas", dp), ("HX", dp), ("cell_edges_Nx", dp), ("cell_edges_Ny", dp),
                ("cell_edges_length", dp), ("cell_stiffness", ip), ("cell_clusters", ip),
                ("node_stencils", dp), ("node2cellsL2up", ip), ("node2cellsL2down", ip), ("cells2nodes", ip)]


class OrcFields(C.Structure):
    _fields_ = [(n, dp) for n in ("densityCells", "UxCells", "UyCells", "pressureCells", "aSoundCells", "VMAXCells",
                                  "densityNodes", "UxNodes", "UyNodes", "pressureNodes")]


class OrcThermo(C.Structure):
    _fields_ = [("RGAS", C.c_double), ("Gamma", C.c_double), ("Cp", C.c_double), ("mu", C.c_double),
                ("kGas", C.c_double)]


class OrcBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("data", C.c_double * 4)]


class OrcParams(C.Structure):
    _fields_ = [("flux", C.c_int32), ("limiter", C.c_int32), ("limiter_eps", C.c_double),
                ("include_udiff", C.c_int32), ("energy_law", C.c_int32), ("mu_const", C.c_double),
                ("k_const", C.c_double), ("bc", OrcBC * (MAX_BC + 1))]


class OrcNsParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu_ref", C.c_double),
                ("T_ref", C.c_double), ("S", C.c_double), ("venkat_K", C.c_double),
                ("venkat_scale", C.c_double * 4), ("flux", C.c_int32),
                ("viscous", C.c_int32), ("mixed", C.c_int32), ("pad", C.c_int32), ("bc", OrcBC * (MAX_BC + 1))]


_lib = None


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_limiter.restype = C.c_double
        L.orc_limiter.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double]
        for fn in (L.orc_ausm_plus_flux, L.orc_ausm_plus_m2020_flux, L.orc_roe_flux):
            fn.restype = None
            fn.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.orc_ns_ausm_flux.restype = None
        L.orc_ns_ausm_flux.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_ns_roe_flux.restype = None
        L.orc_ns_roe_flux.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_ref_bc_ghost.restype = None
        L.orc_ref_bc_ghost.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcThermo), C.POINTER(OrcParams), C.c_int64,
                                       C.c_int32, dp, C.c_double, dp]
        L.orc_air_conductivity.restype = C.c_double
        L.orc_air_conductivity.argtypes = [C.c_double]
        L.orc_sutherland_viscosity.restype = C.c_double
        L.orc_sutherland_viscosity.argtypes = [C.c_double]
        L.orc_distribute_cells_in_threads.argtypes = [C.c_int32, C.c_int64, ip]
        L.orc_phs2dcns.argtypes = [C.POINTER(OrcFields), C.c_int64, C.c_double, dp]
        L.orc_cells2nodes.argtypes = [ip, C.c_int32, C.POINTER(OrcMesh), C.POINTER(OrcFields), dp, dp]
        L.orc_calc_one_stage_muscl_cpu_fast.argtypes = [
            ip, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(OrcMesh), C.POINTER(OrcFields),
            C.POINTER(OrcThermo), C.POINTER(OrcParams), dp, dp, dp, dp]
        L.orc_calc_one_stage_muscl.argtypes = [
            ip, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(OrcMesh), C.POINTER(OrcFields),
            C.POINTER(OrcThermo), C.POINTER(OrcParams), dp, dp, dp, dp, dp]
        L.orc_calc_diff_term.argtypes = [ip, C.c_int32, C.POINTER(OrcMesh), C.POINTER(OrcFields),
                                         C.POINTER(OrcThermo), C.POINTER(OrcParams), dp, dp, dp]
        L.orc_green_gauss_gradient.argtypes = [ip, C.c_int32, C.POINTER(OrcMesh), dp, dp, dp]
        L.orc_update_variables.argtypes = [ip, C.c_int32, C.c_int64, C.c_double, dp, dp, dp, C.POINTER(OrcFields)]
        L.orc_calc_one_stage_cuda.argtypes = [C.c_int32, C.POINTER(OrcMesh), C.POINTER(OrcThermo), C.c_double,
                                              C.c_double, dp, dp, dp, dp, dp]
        L.orc_cuda_split.restype = C.c_double
        L.orc_cuda_split.argtypes = [C.c_int32, C.c_int32, C.c_double]
        L.orc_ref_tau.restype = C.c_double
        L.orc_ref_tau.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcFields), C.c_double]
        L.orc_local_dt.restype = C.c_double
        L.orc_local_dt.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcFields), C.c_double]
        L.orc_ns_step.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcNsParams), C.c_double, dp, dp, dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_ns_step_tl.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcNsParams), C.c_double, dp, dp, dp, dp, dp,
                                     dp, dp]
        L.orc_set_threads.restype = None
        L.orc_ref_loop_tl.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcThermo), C.POINTER(OrcParams), C.c_int32,
                                      C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int32, dp, dp, dp, dp, dp,
                                      dp, dp]
        L.orc_ns_dt.restype = C.c_double
        L.orc_ns_dt.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcNsParams), dp, C.c_double]
        _lib = L
    return _lib


def _p(a, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def _chk_f64(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need float64 column-major"
    return a


class MeshView:
    """Keeps the arrays of a ``mesh2d_Int32`` alive behind an ``OrcMesh``."""

    def __init__(self, mesh):
        self.mesh = mesh
        self._keep = []

        def d(a):
            a = np.asfortranarray(a, dtype=np.float64); self._keep.append(a); return _p(a)

        def i(a):
            a = np.asfortranarray(a, dtype=np.int32); self._keep.append(a); return _p(a, C.c_int32)
        self.c = OrcMesh(mesh.nCells, mesh.nNodes, mesh.nNeibCells, d(mesh.xNodes), d(mesh.yNodes),
                         i(mesh.mesh_connectivity), d(mesh.cell_nodes_X), d(mesh.cell_nodes_Y),
                         d(mesh.cell_mid_points), d(mesh.cell_areas), d(mesh.HX), d(mesh.cell_edges_Nx),
                         d(mesh.cell_edges_Ny), d(mesh.cell_edges_length), i(mesh.cell_stiffness),
                         i(mesh.cell_clusters), d(mesh.node_stencils), i(mesh.node2cellsL2up),
                         i(mesh.node2cellsL2down), i(mesh.cells2nodes))


def fields_struct(f):
    return OrcFields(*[_p(_chk_f64(getattr(f, n))) for n, _ in OrcFields._fields_])


def thermo_struct(th):
    return OrcThermo(th.RGAS, th.Gamma, th.Cp, th.mu, th.kGas)


def bc_table(bcs):
    """``bcs``: dict code -> (kind, data4)."""
    t = (OrcBC * (MAX_BC + 1))()
    for code, (kind, data) in (bcs or {}).items():
        idx = -int(code)
        t[idx].kind = kind
        for q, v in enumerate(list(data) + [0.0] * (4 - len(data))):
            t[idx].data[q] = v
    return t


def params(flux=FLUX_AUSM_PLUS, limiter=LIM_REF_HARMONIC, limiter_eps=1e-12, include_udiff=True,
           energy_law=ENERGY_AIR_POLY, mu_const=1e-2, k_const=0.0, bcs=None):
    return OrcParams(flux, limiter, limiter_eps, int(include_udiff), energy_law, mu_const, k_const, bc_table(bcs))


def ref_bc_ghost(mv: MeshView, th, prm, i, k, cur, t=0.0):
    """The reference scheme's ghost primitives of face k of cell i (both
    1-based) from the cell's primitives ``cur`` at time ``t``."""
    c = np.ascontiguousarray(cur, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_ref_bc_ghost(C.byref(mv.c), C.byref(thermo_struct(th)), C.byref(prm), int(i), int(k), _p(c), float(t),
                           _p(out))
    return out


def threads_table(n_threads, n):
    t = np.zeros((n_threads, 2), dtype=np.int32, order="F")
    lib().orc_distribute_cells_in_threads(n_threads, n, _p(t, C.c_int32))
    return t


# ---------------------------------------------------------------- pointwise

def limiter(kind, x, y, c=1e-12):
    return lib().orc_limiter(kind, x, y, c)


def _flux(fn, qL, qR, nx, ny, side, gamma):
    q = np.ascontiguousarray(np.concatenate([qL, qR]), dtype=np.float64)
    out = np.zeros(4)
    fn(_p(q), nx, ny, side, gamma, _p(out))
    return out


def ausm_plus_flux(qL, qR, nx, ny, side=1.0, gamma=1.4):
    return _flux(lib().orc_ausm_plus_flux, qL, qR, nx, ny, side, gamma)


def ausm_plus_m2020_flux(qL, qR, nx, ny, side=1.0, gamma=1.4):
    return _flux(lib().orc_ausm_plus_m2020_flux, qL, qR, nx, ny, side, gamma)


def roe_flux(qL, qR, nx, ny, side=1.0, gamma=1.4):
    return _flux(lib().orc_roe_flux, qL, qR, nx, ny, side, gamma)


def _ns_flux(fn, qL, qR, nx, ny, gamma):
    q = np.ascontiguousarray(np.concatenate([qL, qR]), dtype=np.float64)
    out = np.zeros(4)
    fn(_p(q), nx, ny, gamma, _p(out))
    return out


def ns_ausm_flux(qL, qR, nx, ny, gamma=1.4):
    """The ns scheme's AUSM+ per unit length (qL: the neighbour side)."""
    return _ns_flux(lib().orc_ns_ausm_flux, qL, qR, nx, ny, gamma)


def ns_roe_flux(qL, qR, nx, ny, gamma=1.4):
    """The ns scheme's Roe per unit length (qL: the neighbour side)."""
    return _ns_flux(lib().orc_ns_roe_flux, qL, qR, nx, ny, gamma)


def air_conductivity(T):
    return lib().orc_air_conductivity(T)


def sutherland_viscosity(T):
    return lib().orc_sutherland_viscosity(T)


# ---------------------------------------------------------------- operators

def _rc(rc, name):
    if rc != 0:
        raise RuntimeError(f"oracle {name} failed with status {rc}")


def phs2dcns(fields, gamma, n):
    U = np.zeros((n, 4), order="F")
    f = fields_struct(fields)
    lib().orc_phs2dcns(C.byref(f), n, gamma, _p(U))
    return U


def cells2nodes(mv: MeshView, fields, Ucells=None, n_threads=2):
    """Node primitives (in ``fields``, in place) and node conservative values."""
    m = mv.mesh
    t = threads_table(n_threads, m.nNodes)
    Un = np.zeros((m.nNodes, 4), order="F") if Ucells is not None else None
    f = fields_struct(fields)
    _rc(lib().orc_cells2nodes(_p(t, C.c_int32), n_threads, C.byref(mv.c), C.byref(f),
                              _p(None if Ucells is None else _chk_f64(Ucells)), _p(Un)), "cells2nodes")
    return Un


def calc_one_stage_muscl_cpu_fast(mv: MeshView, fields, thermo, prm, UOld, Udiff, betta, dt, flow_time,
                                  n_threads=2, threads=None):
    m = mv.mesh
    t = threads if threads is not None else threads_table(n_threads, m.nCells)
    FL = np.zeros((m.nCells, 4), order="F")
    UN = np.zeros((m.nCells, 4), order="F")
    f = fields_struct(fields); th = thermo_struct(thermo)
    _rc(lib().orc_calc_one_stage_muscl_cpu_fast(_p(t, C.c_int32), t.shape[0], betta, dt, flow_time, C.byref(mv.c),
                                                C.byref(f), C.byref(th), C.byref(prm), _p(_chk_f64(UOld)), _p(FL),
                                                _p(None if Udiff is None else _chk_f64(Udiff)), _p(UN)), "step")
    return FL, UN


def calc_one_stage_muscl(mv: MeshView, fields, thermo, prm, UOld, betta, dt, flow_time, n_threads=2,
                         face_states=False):
    """calcOneStageMUSCL; with ``face_states`` also returns its phase-1
    arrays ULeft1..4 and URight1..4, each ``(4, nCells)`` column-major, as
    ``(4 slots, 4, nCells)`` numpy arrays."""
    m = mv.mesh
    t = threads_table(n_threads, m.nCells)
    FL = np.zeros((m.nCells, 4), order="F")
    UN = np.zeros((m.nCells, 4), order="F")
    UL = np.zeros(16 * m.nCells) if face_states else None
    UR = np.zeros(16 * m.nCells) if face_states else None
    f = fields_struct(fields); th = thermo_struct(thermo)
    _rc(lib().orc_calc_one_stage_muscl(_p(t, C.c_int32), n_threads, betta, dt, flow_time, C.byref(mv.c), C.byref(f),
                                       C.byref(th), C.byref(prm), _p(_chk_f64(UOld)), _p(UN), _p(FL), _p(UL), _p(UR)),
        "two-phase step")
    if not face_states:
        return FL, UN
    # slot k's (4, n) column-major array: element [v, i] at 4 i + v
    shape = lambda a: a.reshape(4, m.nCells, 4).transpose(0, 2, 1)  # noqa: E731
    return FL, UN, shape(UL), shape(UR)


def calc_one_stage_cuda(mv: MeshView, thermo, betta, dt, UOld, ULeft, URight, x32=False):
    """calcOneStageCUDA / calcOneStageCUDAx32 from the face states ``ULeft``,
    ``URight`` ((4 slots, 4, nCells) as calc_one_stage_muscl returns them):
    returns (iFlux (nCells, 4), UNew)."""
    m = mv.mesh
    raw = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(0, 2, 1))  # noqa: E731
    UL, UR = raw(ULeft), raw(URight)
    FL = np.zeros((m.nCells, 4), order="F")
    UN = np.zeros((m.nCells, 4), order="F")
    th = thermo_struct(thermo)
    _rc(lib().orc_calc_one_stage_cuda(int(bool(x32)), C.byref(mv.c), C.byref(th), betta, dt, _p(_chk_f64(UOld)),
                                      _p(UL), _p(UR), _p(UN), _p(FL)), "calcOneStageCUDA")
    return FL, UN


def cuda_split(x32, which, M):
    """The GPU kernels' AUSM+ split functions (0 M+, 1 M-, 2 P+, 3 P-)."""
    return lib().orc_cuda_split(int(bool(x32)), int(which), float(M))


def calc_diff_term(mv: MeshView, fields, thermo, prm, UC, Udiff_in=None, n_threads=2):
    m = mv.mesh
    t = threads_table(n_threads, m.nCells)
    Ud = np.zeros((m.nCells, 4), order="F") if Udiff_in is None else np.array(Udiff_in, order="F", copy=True)
    lap = np.zeros((m.nCells, 3), order="F")
    f = fields_struct(fields); th = thermo_struct(thermo)
    _rc(lib().orc_calc_diff_term(_p(t, C.c_int32), n_threads, C.byref(mv.c), C.byref(f), C.byref(th), C.byref(prm),
                                 _p(_chk_f64(UC)), _p(Ud), _p(lap)), "calcDiffTerm")
    return Ud, lap


def green_gauss_gradient(mv: MeshView, phi, n_threads=2):
    m = mv.mesh
    t = threads_table(n_threads, m.nCells)
    gx = np.zeros(m.nCells); gy = np.zeros(m.nCells)
    _rc(lib().orc_green_gauss_gradient(_p(t, C.c_int32), n_threads, C.byref(mv.c),
                                       _p(np.ascontiguousarray(phi, dtype=np.float64)), _p(gx), _p(gy)), "GG")
    return gx, gy


def update_variables(fields, gamma, UNew, UOld, n_threads=2):
    """updateVariablesSA: fields and UOld are updated in place; returns Delta."""
    n = UNew.shape[0]
    t = threads_table(n_threads, n)
    Delta = np.zeros((n, 4), order="F")
    f = fields_struct(fields)
    _rc(lib().orc_update_variables(_p(t, C.c_int32), n_threads, n, gamma, _p(_chk_f64(UNew)), _p(_chk_f64(UOld)),
                                   _p(Delta), C.byref(f)), "updateVariablesSA")
    return Delta


def ref_tau(mv: MeshView, fields, cfl):
    f = fields_struct(fields)
    return lib().orc_ref_tau(C.byref(mv.c), C.byref(f), cfl)


def local_dt(mv: MeshView, fields, cfl):
    f = fields_struct(fields)
    return lib().orc_local_dt(C.byref(mv.c), C.byref(f), cfl)


def ns_params(gamma=1.4, R=287.058, Pr=0.72, mu_ref=1.8476e-5, T_ref=300.0, S=116.0, venkat_K=5.0,
              venkat_scale=(1.0, 1.0, 1.0, 1.0), flux=FLUX_AUSM_PLUS, viscous=True, mixed=False, bcs=None):
    p = OrcNsParams()
    p.gamma, p.R, p.Pr, p.mu_ref, p.T_ref, p.S, p.venkat_K = gamma, R, Pr, mu_ref, T_ref, S, venkat_K
    for q in range(4):
        p.venkat_scale[q] = venkat_scale[q]
    p.flux, p.viscous, p.mixed = flux, int(viscous), int(mixed)
    t = bc_table(bcs)
    for q in range(MAX_BC + 1):
        p.bc[q] = t[q]
    return p


def ns_step(mv: MeshView, prm, dt, U, want_grad=False):
    """One ns-scheme step; returns (U_new, grad8 or None, phi4 or None, wall force (fx, fy))."""
    n = mv.mesh.nCells
    UN = np.zeros((n, 4), order="F")
    g8 = np.zeros((n, 8), order="F") if want_grad else None
    p4 = np.zeros((n, 4), order="F") if want_grad else None
    f2 = np.zeros(2)
    _rc(lib().orc_ns_step(C.byref(mv.c), C.byref(prm), dt, _p(_chk_f64(U)), _p(UN), _p(g8), _p(p4), _p(f2)), "ns")
    return UN, g8, p4, f2


def ns_step_tl(mv: MeshView, prm, dt, U, dU, dbc=None):
    """Tangent-linear FP64 ns step (dual numbers): returns (U_new, dU_new, force2, dforce2).
    dbc: tangent of the fixed boundary states, shape (MAX_BC + 1, 4), or None."""
    n = mv.mesh.nCells
    UN = np.zeros((n, 4), order="F")
    dUN = np.zeros((n, 4), order="F")
    f2, df2 = np.zeros(2), np.zeros(2)
    db = None if dbc is None else np.ascontiguousarray(dbc, dtype=np.float64).reshape(-1)
    if db is not None and db.size != 4 * (MAX_BC + 1):
        raise ValueError("dbc must have (MAX_BC + 1) x 4 entries")
    _rc(lib().orc_ns_step_tl(C.byref(mv.c), C.byref(prm), dt, _p(_chk_f64(U)), _p(_chk_f64(dU)), _p(db), _p(UN),
                             _p(dUN), _p(f2), _p(df2)), "ns_tl")
    return UN, dUN, f2, df2


def ref_loop_tl(mv: MeshView, thermo, prm, U0, dU0, dt, n_steps, viscous=True, betta=1.0, flow_time=0.0):
    """N steps of the reference loop (calcDiffTerm, calcOneStageMUSCLcpuFast,
    updateVariablesSA, node interpolation) from U0 at a fixed dt, starting from
    the primitives of U0 and their node means, the flow time advancing from
    flow_time by dt per step, with the tangent dU0 carried in dual numbers:
    returns (U_N, dU_N).  AUSM+, constant mu and k; triangles and quads,
    periodic or with the boundary ghosts."""
    UN, dUN, _, _ = ref_loop_tl_force(mv, thermo, prm, U0, dU0, dt, n_steps, 0, viscous, betta, flow_time)
    return UN, dUN


def ref_loop_tl_force(mv: MeshView, thermo, prm, U0, dU0, dt, n_steps, wall_code, viscous=True, betta=1.0,
                      flow_time=0.0, dbc=None):
    """ref_loop_tl with the wall force of the faces of code wall_code (the
    AUSM+ momentum flux through them) summed over the N steps, and the
    tangent dbc ((MAX_BC + 1, 4) by -code, or None) of the fixed states:
    returns (U_N, dU_N, force (2,), dforce (2,))."""
    n = mv.mesh.nCells
    UN = np.zeros((n, 4), order="F")
    dUN = np.zeros((n, 4), order="F")
    F = np.zeros(2)
    dF = np.zeros(2)
    th = thermo_struct(thermo)
    d = None if dU0 is None else _chk_f64(dU0)
    b = None if dbc is None else np.ascontiguousarray(dbc, dtype=np.float64).reshape(MAX_BC + 1, 4)
    _rc(lib().orc_ref_loop_tl(C.byref(mv.c), C.byref(th), C.byref(prm), int(bool(viscous)), betta, dt, int(n_steps),
                              float(flow_time), int(wall_code), _p(_chk_f64(U0)), _p(d), _p(b), _p(UN), _p(dUN),
                              _p(F), _p(dF)), "ref_loop_tl")
    return UN, dUN, F, dF


def ns_dt(mv: MeshView, prm, U, cfl):
    return lib().orc_ns_dt(C.byref(mv.c), C.byref(prm), _p(_chk_f64(U)), cfl)


def set_threads(n: int):
    lib().orc_set_threads(int(n))
