"""Host-side mirror of the reference's step interface, backed by libgodunov.

The functions keep the reference's names, positional arguments and in-place
semantics, so a time loop written against the reference reads the same:

* ``calcOneStageMUSCLcpuFast``  calcOneStageMUSCLcpuFast.jl:3-107
* ``calcOneStageMUSCL``         calcOneStageMUSCL.jl:4-106
* ``calcOneStageCUDA`` / ``calcOneStageCUDAx32``  calcOneStageCUDA.jl:4-135, calcOneStageCUDAx32.jl:5-155
* ``calcDiffTerm``              calcDiffTerm.jl:31-121
* ``updateVariablesSA``         evaluate2d.jl:52-110
* ``cells2nodesSolutionReconstructionWithStencils``  utilsFVM2dp.jl:299-365
* ``phs2dcns2dcellsSA``         utilsFVM2dp.jl:4-17

Each call uploads its inputs to HBM, runs the HIP kernels and copies the
outputs back into the caller's arrays (column-major ``(nCells, 4)`` float64,
1-based ids in the mesh).  ``GodunovSolver`` keeps everything resident in HBM
for a whole run (the performance path).

The reference chooses its numerics by which files it ``include``s; here the
same choices live in ``SchemeConfig`` (``configure()`` sets the module
default).  The default is the reference's production set: AUSM+ fluxes, the
harmonic ``Minmod_Limiter`` with eps 1e-12, the generic boundary conditions of
boundaryConditions2d.jl (code -1 fixed M=3.55 inflow, -2 zero-gradient,
-3 slip wall) and calcDiffTerm's mu = 1e-2 with the air k(T).
"""
from __future__ import annotations

import dataclasses
import weakref
from typing import Dict, Optional, Tuple

import numpy as np

from . import _lib as L

REFERENCE_BCS = {
    -1: (L.GD_BC_FIXED, (1.1766, 1232.6445, 0.0, 101325.0)),  # boundaryConditions2d.jl:34-39
    -2: (L.GD_BC_EXTRAPOLATE, ()),                              # boundaryConditions2d.jl:47-52
    -3: (L.GD_BC_SLIP_WALL, ()),                                # boundaryConditions2d.jl:41-45
}

TGV_BCS = {code: (L.GD_BC_TGV_MIRROR, (100000.0, 40.0, 1.0, 1e-2)) for code in range(0, -L.GD_MAX_BC - 1, -1)}
"""boundaryConditions_TGV2d.jl:16-30 with computeTGV2d's P, U, L, mu."""


@dataclasses.dataclass
class SchemeConfig:
    flux: int = L.GD_FLUX_AUSM_PLUS
    limiter: int = L.GD_LIM_REF_HARMONIC
    limiter_eps: float = 1.0e-12
    bcs: Dict[int, Tuple[int, tuple]] = dataclasses.field(default_factory=lambda: dict(REFERENCE_BCS))
    energy_law: int = L.GD_ENERGY_AIR_POLY
    mu_const: float = 1.0e-2
    k_const: float = 0.0

    def params(self, thermo) -> L.gd_params:
        p = L.default_params()
        p.scheme = L.GD_SCHEME_REF
        p.precision = L.GD_FP64
        p.flux = self.flux
        p.limiter = self.limiter
        p.limiter_eps = self.limiter_eps
        p.viscous = 1
        p.energy_law = self.energy_law
        p.mu_const = self.mu_const
        p.k_const = self.k_const
        p.RGAS = thermo.RGAS
        p.Gamma = thermo.Gamma
        L.set_bcs(p, self.bcs)
        return p


_config = SchemeConfig()


def configure(**kw) -> SchemeConfig:
    """Set the module's default numerics (see ``SchemeConfig``)."""
    global _config
    _config = dataclasses.replace(_config, **kw)
    return _config


class _Context:
    def __init__(self, mesh):
        self.dmesh = L.DeviceMesh(mesh)
        self.state = None

    def state_for(self, params):
        if self.state is None:
            self.state = L.DeviceState(self.dmesh, params)
        else:
            self.state.set_params(params)
        return self.state


_contexts: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()
_contexts_by_id: Dict[int, Tuple[object, _Context]] = {}


def _ctx(mesh) -> _Context:
    try:
        c = _contexts.get(mesh)
    except TypeError:
        c = None
    if c is None:
        ent = _contexts_by_id.get(id(mesh))
        if ent is not None and ent[0] is mesh:
            return ent[1]
        c = _Context(mesh)
        _UPDATE_MESH[mesh.nCells] = mesh
        try:
            _contexts[mesh] = c
        except TypeError:
            _contexts_by_id[id(mesh)] = (mesh, c)
    return c


def _ranges(threads, n):
    """Cells covered by a reference partition table (1-based inclusive rows).

    A row of zeros is the one-thread table of distibuteCellsInThreadsSA; the
    reference then indexes cell 0 and throws, so this raises IndexError."""
    t = np.asarray(threads)
    mask = np.zeros(n, dtype=bool)
    for b, e in t.reshape(-1, 2):
        b, e = int(b), int(e)
        if b < 1 or e > n:
            raise IndexError(f"partition row {b}:{e} indexes outside 1:{n} (BoundsError in the reference)")
        mask[b - 1:e] = True
    return mask


def _set_cells(st: L.DeviceState, f, nodes=True):
    if nodes:
        st.set_fields(f.densityCells, f.UxCells, f.UyCells, f.pressureCells,
                      f.densityNodes, f.UxNodes, f.UyNodes, f.pressureNodes)
    else:
        st.set_fields(f.densityCells, f.UxCells, f.UyCells, f.pressureCells)


def _copy_rows(dst, src, mask):
    if mask.all():
        dst[...] = src
    else:
        dst[mask] = src[mask]


def calcOneStageMUSCLcpuFast(cellsThreads, bettaKJ, dt, flowTime, testMesh, testFields, thermo,  # noqa: N802
                             UConsCellsOld, FLUXES, UconsDiffTerm, UconsCellsNew, *, config: SchemeConfig = None):
    """One explicit stage with the viscous term: ``FLUXES`` and ``UconsCellsNew`` are overwritten."""
    mask = _ranges(cellsThreads, testMesh.nCells)
    st = _ctx(testMesh).state_for((config or _config).params(thermo))
    _set_cells(st, testFields)
    st.set_array(L.GD_ARR_U_OLD, UConsCellsOld)
    st.set_array(L.GD_ARR_U_DIFF, UconsDiffTerm)
    st.calc_one_stage(L.GD_STAGE_MUSCL_FAST, bettaKJ, dt, flowTime)
    _copy_rows(FLUXES, st.get_array(L.GD_ARR_FLUXES), mask)
    _copy_rows(UconsCellsNew, st.get_array(L.GD_ARR_U_NEW), mask)


def calcOneStageMUSCL(betta, dtX, flowTime, testMesh, testfields2dX, thermo, cellsThreadsX,  # noqa: N802
                      UconsCellsOldX, UconsDiffTermX, UconsCellsNewX, ULeft1=None, ULeft2=None, ULeft3=None,
                      ULeft4=None, URight1=None, URight2=None, URight3=None, URight4=None, FLUXES=None, *,
                      config: SchemeConfig = None):
    """Two-phase stage without the viscous term (the source comments it out).

    The update itself runs as one fused kernel (face states in registers);
    when the caller passes the face-state arrays ``ULeft1..4`` /
    ``URight1..4`` (each ``(4, nCells)``), they are filled as the reference's
    phase one fills them (computeMUSCLStencilsCUDA.jl:3-28): every slot of a
    cell's faces, slot 4 of triangles left as given."""
    mask = _ranges(cellsThreadsX, testMesh.nCells)
    st = _ctx(testMesh).state_for((config or _config).params(thermo))
    _set_cells(st, testfields2dX)
    st.set_array(L.GD_ARR_U_OLD, UconsCellsOldX)
    st.calc_one_stage(L.GD_STAGE_MUSCL, betta, dtX, flowTime)
    _copy_rows(UconsCellsNewX, st.get_array(L.GD_ARR_U_NEW), mask)
    if FLUXES is not None:
        _copy_rows(FLUXES, st.get_array(L.GD_ARR_FLUXES), mask)
    lefts, rights = (ULeft1, ULeft2, ULeft3, ULeft4), (URight1, URight2, URight3, URight4)
    if any(a is not None for a in lefts + rights):
        nc = testMesh.nCells
        cur = lambda arrs: np.stack([np.zeros((4, nc)) if a is None else np.asarray(a, dtype=np.float64)  # noqa: E731
                                     for a in arrs])
        UL, UR = st.calc_face_states(flowTime, cur(lefts), cur(rights))
        for arrs, full in ((lefts, UL), (rights, UR)):
            for k, a in enumerate(arrs):
                if a is not None:
                    a[:, mask] = full[k][:, mask]


def _cuda_stage(x32, betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX, UconsCellsOldX,
                UconsCellsNewX, lefts, rights, fluxes, config):
    mask = _ranges(cellsThreadsX, testMeshX.nCells)
    st = _ctx(testMeshX).state_for((config or _config).params(thermoX))
    st.set_array(L.GD_ARR_U_OLD, UconsCellsOldX)
    nc = testMeshX.nCells
    have = all(a is not None for a in lefts[:3] + rights[:3])
    if not have:  # no face states handed in: phase one on the device from the fields
        _set_cells(st, testfields2dX)
    UL = np.stack([np.asarray(a, dtype=np.float64) if a is not None else np.zeros((4, nc)) for a in lefts])
    UR = np.stack([np.asarray(a, dtype=np.float64) if a is not None else np.zeros((4, nc)) for a in rights])
    iflux = st.calc_one_stage_cuda(L.GD_FP32_MIXED if x32 else L.GD_FP64, betta, dtX, flowTime,
                                   UL if have else None, UR if have else None)
    _copy_rows(UconsCellsNewX, st.get_array(L.GD_ARR_U_NEW), mask)
    # copyto!(iFluxV_k, cuFluxV_k) copies every cell (calcOneStageCUDAx32.jl:126-129)
    for k, a in enumerate(fluxes):
        if a is not None:
            a[:] = iflux[:, k]


def calcOneStageCUDA(betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX,  # noqa: N802
                     UconsCellsOldX, UconsDiffTermX, UconsCellsNewX, uLeft1=None, uLeft2=None, uLeft3=None,
                     uLeft4=None, uRight1=None, uRight2=None, uRight3=None, uRight4=None, iFluxV1=None,
                     iFluxV2=None, iFluxV3=None, iFluxV4=None, *device_buffers, config: SchemeConfig = None):
    """The reference's FP64 GPU stage: AUSM+ of the given face states
    (computeMUSCLStencilsCUDA's uLeft/uRight) on face slots 1..3, summed per
    cell on the device, then the update without the viscous term.  The CuVector
    work buffers of the source are not needed (the library owns its device
    memory); without face states, phase one runs on the device."""
    _cuda_stage(False, betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX, UconsCellsOldX,
                UconsCellsNewX, (uLeft1, uLeft2, uLeft3, uLeft4), (uRight1, uRight2, uRight3, uRight4),
                (iFluxV1, iFluxV2, iFluxV3, iFluxV4), config)


def calcOneStageCUDAx32(betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX,  # noqa: N802
                        UconsCellsOldX, UconsDiffTermX, UconsCellsNewX, uLeft1=None, uLeft2=None, uLeft3=None,
                        uLeft4=None, uRight1=None, uRight2=None, uRight3=None, uRight4=None, iFluxV1=None,
                        iFluxV2=None, iFluxV3=None, iFluxV4=None, *device_buffers, config: SchemeConfig = None):
    """The reference's mixed GPU stage: the face states rounded to Float32,
    AUSM+ and the per-cell sums in Float32 (with the Float64 literals of its
    split functions), the update in Float64 without the viscous term."""
    _cuda_stage(True, betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX, UconsCellsOldX,
                UconsCellsNewX, (uLeft1, uLeft2, uLeft3, uLeft4), (uRight1, uRight2, uRight3, uRight4),
                (iFluxV1, iFluxV2, iFluxV3, iFluxV4), config)


def calcDiffTerm(cellsThreads, nodesThreads, testMesh, testfields2d, viscousFields2d, thermo,  # noqa: N802
                 UCCells, UCNodes, UConsDiffCells, *, config: SchemeConfig = None):
    """Viscous term: overwrites columns 2:4 of ``UConsDiffCells`` (column 1 is left as is)."""
    mask = _ranges(cellsThreads, testMesh.nCells)
    st = _ctx(testMesh).state_for((config or _config).params(thermo))
    _set_cells(st, testfields2d, nodes=False)
    st.set_array(L.GD_ARR_U_OLD, UCCells)
    st.set_array(L.GD_ARR_U_DIFF, UConsDiffCells)
    st.calc_diff_term()
    _copy_rows(UConsDiffCells, st.get_array(L.GD_ARR_U_DIFF), mask)


def updateVariablesSA(cellsThreadsX, Gamma, UconsCellsNew, UconsCellsOld, Delta, testfields2d,  # noqa: N802
                      solControls=None, dynControls=None):
    """Primitive recovery from ``UconsCellsNew``; ``Delta = new - old``; ``UconsCellsOld = new``."""
    n = UconsCellsNew.shape[0]
    mask = _ranges(cellsThreadsX, n)
    if dynControls is not None:
        dynControls.rhoMax = float(np.max(testfields2d.densityCells))
        dynControls.rhoMin = float(np.min(testfields2d.densityCells))
    mesh = _UPDATE_MESH.get(n)
    if mesh is None:
        raise RuntimeError("updateVariablesSA needs a mesh context: call a step operator or register_mesh first")
    st = _ctx(mesh).state
    p = st.params
    p.Gamma = Gamma
    st.set_params(p)
    st.set_array(L.GD_ARR_U_NEW, UconsCellsNew)
    st.set_array(L.GD_ARR_U_OLD, UconsCellsOld)
    st.update_variables()
    rC, uC, vC, pC, aC, vmC = st.get_fields()[:6]
    for name, arr in (("densityCells", rC), ("UxCells", uC), ("UyCells", vC), ("pressureCells", pC),
                      ("aSoundCells", aC), ("VMAXCells", vmC)):
        dst = getattr(testfields2d, name)
        dst[mask] = arr[mask]
    _copy_rows(Delta, st.get_array(L.GD_ARR_DELTA), mask)
    _copy_rows(UconsCellsOld, st.get_array(L.GD_ARR_U_OLD), mask)


_UPDATE_MESH: Dict[int, object] = {}


def register_mesh(testMesh):
    """Associates arrays of ``testMesh.nCells`` rows with this mesh for
    ``updateVariablesSA`` (whose reference signature carries no mesh)."""
    _UPDATE_MESH[testMesh.nCells] = testMesh
    _ctx(testMesh)


def cells2nodesSolutionReconstructionWithStencils(nodesThreads, testMesh, testfields2d, viscfields2d,  # noqa: N802
                                                  cell_solution, node_solution, *, config: SchemeConfig = None,
                                                  thermo=None):
    """Node primitives (into ``testfields2d``) and node conservative values (into ``node_solution``)."""
    _ranges(nodesThreads, testMesh.nNodes)
    ctx = _ctx(testMesh)
    st = ctx.state
    if st is None:
        from .mesh import THERMOPHYSICS
        st = ctx.state_for((config or _config).params(thermo or THERMOPHYSICS()))
    _set_cells(st, testfields2d, nodes=False)
    st.set_array(L.GD_ARR_U_OLD, cell_solution)
    st.cells2nodes()
    f = st.get_fields()
    testfields2d.densityNodes[:] = f[6]
    testfields2d.UxNodes[:] = f[7]
    testfields2d.UyNodes[:] = f[8]
    testfields2d.pressureNodes[:] = f[9]
    node_solution[...] = st.get_array(L.GD_ARR_U_NODES)


def phs2dcns2dcellsSA(ACons, testFields, gamma):  # noqa: N802
    """Conservative from primitive cell values, utilsFVM2dp.jl:4-17 (host arithmetic)."""
    r, u, v, p = testFields.densityCells, testFields.UxCells, testFields.UyCells, testFields.pressureCells
    ACons[:, 0] = r
    ACons[:, 1] = r * u
    ACons[:, 2] = r * v
    ACons[:, 3] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)


# ---------------------------------------------------------------------------
# device-resident time loop
# ---------------------------------------------------------------------------

def make_params(scheme=L.GD_SCHEME_REF, precision=L.GD_FP64, flux=L.GD_FLUX_AUSM_PLUS, limiter=None,
                limiter_eps=1e-12, venkat_K=5.0, venkat_scale=(1.0, 1.0, 1.0, 1.0), viscous=True, energy_law=L.GD_ENERGY_AIR_POLY, mu_const=1e-2,
                k_const=0.0, RGAS=287.058, Gamma=1.4, mu_ref=1.8476e-5, T_ref=300.0, S_suth=116.0, Pr=0.72,
                bcs=None) -> L.gd_params:
    p = L.default_params()
    p.scheme = scheme
    p.precision = precision
    p.flux = flux
    p.limiter = (L.GD_LIM_REF_HARMONIC if scheme == L.GD_SCHEME_REF else L.GD_LIM_VENKATAKRISHNAN) \
        if limiter is None else limiter
    p.limiter_eps = limiter_eps
    p.venkat_K = venkat_K
    for q in range(4):
        p.venkat_scale[q] = float(venkat_scale[q])
    p.viscous = int(bool(viscous))
    p.energy_law = energy_law
    p.mu_const = mu_const
    p.k_const = k_const
    p.RGAS = RGAS
    p.Gamma = Gamma
    p.mu_ref = mu_ref
    p.T_ref = T_ref
    p.S_suth = S_suth
    p.Pr = Pr
    L.set_bcs(p, REFERENCE_BCS if bcs is None else bcs)
    return p


class GodunovSolver:
    """A mesh and a state resident in HBM, advanced by ``gd_run``."""

    def __init__(self, mesh, params: L.gd_params, order: int = L.GD_ORDER_AS_GIVEN, perm=None):
        self.mesh = mesh
        self.dmesh = L.DeviceMesh(mesh, order, perm)
        self.state = L.DeviceState(self.dmesh, params)
        self.flow_time = 0.0

    @property
    def params(self):
        return self.state.params

    def set_primitives(self, rho, u, v, p, nodes=None):
        """Cell primitives; node primitives from ``nodes`` = (rho,u,v,p) or, if
        None, by the node interpolation of the same step (the reference's
        drivers start with zero node fields, which this avoids)."""
        self.state.set_fields(rho, u, v, p, *(nodes or (None,) * 4))
        U = np.zeros((len(rho), 4), order="F")
        g = self.state.params.Gamma
        U[:, 0] = rho; U[:, 1] = rho * u; U[:, 2] = rho * v
        U[:, 3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
        self.state.set_array(L.GD_ARR_U_OLD, U)
        self.state.set_array(L.GD_ARR_U_NEW, U)
        if nodes is None:
            self.state.cells2nodes()

    def set_conservative(self, U):
        """Sets U_OLD and the primitives recovered from it (updateVariablesSA), then node values."""
        self.state.set_array(L.GD_ARR_U_NEW, U)
        self.state.update_variables()
        self.state.cells2nodes()

    def run(self, n_steps, dt=None, cfl=None, dt_policy=None, betta=1.0, time_kernels=False, residuals=False,
            wall_code=0, raise_nonphysical=True):
        rp = L.gd_run_params()
        if dt_policy is None:
            dt_policy = L.GD_DT_FIXED if dt is not None else L.GD_DT_LOCAL
        rp.dt_policy = dt_policy
        rp.dt = float(dt or 0.0)
        rp.cfl = float(cfl or 0.0)
        rp.betta = float(betta)
        rp.flow_time = float(self.flow_time)
        rp.time_kernels = int(time_kernels)
        rp.residuals = int(residuals)
        rp.wall_code = int(wall_code)
        st = self.state.run(int(n_steps), rp, raise_nonphysical=raise_nonphysical)
        self.flow_time = rp.flow_time
        return st

    def adjoint(self, n_steps, dt, checkpoint_every=20, wall_code=0, force_weight=(1.0, 0.0), dJ_dUN=None,
                record_budget=0, to_host=True):
        """Discrete adjoint of ``n_steps`` fixed-dt steps of the solver's
        scheme from the current state (which ends at U_N, as after ``run``)
        for J = sum_n w . F_n + <dJ_dUN, U_N>, F_n the wall force of step n on
        ``wall_code`` (the reference scheme: AUSM+ with an odd, symmetric
        limiter, constant mu and k).  ``record_budget`` (bytes) keeps the cell records of as
        many trailing checkpoint segments as fit instead of recomputing them.
        ``to_host=False`` leaves dJ/dU0 on the device (gd_state_device_adjoint).
        Returns (dJ/dU0 (n_cells, 4) or None, dJ/dbc (MAX_BC + 1, 4),
        gd_adjoint_stats)."""
        ap = L.gd_adjoint_params()
        ap.n_steps = int(n_steps)
        ap.checkpoint_every = int(checkpoint_every)
        ap.dt = float(dt)
        ap.flow_time = float(self.flow_time)
        ap.wall_code = int(wall_code)
        ap.force_weight[0], ap.force_weight[1] = float(force_weight[0]), float(force_weight[1])
        ap.record_budget = int(record_budget)
        g0, gbc, st = self.state.adjoint(ap, dJ_dUN, to_host)
        self.flow_time = st.flow_time
        return g0, gbc, st

    def conservative(self):
        return self.state.get_array(L.GD_ARR_U_OLD)

    def fields(self):
        return self.state.get_fields()

    def save_snapshot(self, path, dt=0.0):
        """Cell primitives and flow time to an HDF5 restart file
        (``solutionCellsT`` fields)."""
        from .hdf5io import save_solution_hdf5
        f = self.state.get_fields()
        save_solution_hdf5(path, dt, self.flow_time, f[0], f[1], f[2], f[3])

    def load_snapshot(self, path):
        """Restart from ``save_snapshot`` (as loadPrevResults.jl:3-38): cell
        primitives, derived node values and the flow time."""
        from .hdf5io import load_solution_hdf5
        d = load_solution_hdf5(path)
        if d["nCells"] != self.mesh.nCells:
            raise ValueError(f"snapshot has {d['nCells']} cells, the mesh {self.mesh.nCells}")
        self.set_primitives(d["densityCells"], d["UxCells"], d["UyCells"], d["pressureCells"])
        self.flow_time = d["flowTime"]
        return d

    def save_vtk(self, filename, binary=True):
        """The node primitives (the node interpolation of the current cell
        state) as a .vtu file, like saveResults4VTK (utilsIO.jl:43-53)."""
        from .vtkio import save_results_vtk
        self.state.cells2nodes()
        f = self.state.get_fields()
        return save_results_vtk(filename, self.mesh, {"density": f[6], "Ux": f[7], "Uy": f[8], "pressure": f[9]},
                                binary=binary)

    def close(self):
        self.state.close()
        self.dmesh.close()


# ---------------------------------------------------------------------------
# supersonic cylinder configuration: free stream and gas
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class CylinderCase:
    """M = 3.5 air flow over a cylinder of diameter 1 (mesh units) at Re_D.

    Dimensional air at T = 300 K, p = 101325 Pa; the Sutherland law has the
    reference's constants (sutherland.jl: As = 1.4792e-6, Ts = 116 K) with its
    scale chosen so that mu(T_inf) gives Re_D on a diameter-1 cylinder."""
    mach: float = 3.5
    re_d: float = 1.0e5
    T_inf: float = 300.0
    p_inf: float = 101325.0
    T_wall: float = 300.0
    gamma: float = 1.4
    R: float = 287.058
    Pr: float = 0.72
    S: float = 116.0

    @property
    def rho_inf(self):
        return self.p_inf / (self.R * self.T_inf)

    @property
    def u_inf(self):
        return self.mach * (self.gamma * self.R * self.T_inf) ** 0.5

    @property
    def mu_inf(self):
        return self.rho_inf * self.u_inf * 1.0 / self.re_d

    def bcs(self):
        return {-1: (L.GD_BC_FIXED, (self.rho_inf, self.u_inf, 0.0, self.p_inf)),
                -2: (L.GD_BC_EXTRAPOLATE, ()),
                -3: (L.GD_BC_ISOTHERMAL_WALL, (self.T_wall,))}

    def params(self, precision=L.GD_FP64, flux=L.GD_FLUX_AUSM_PLUS, viscous=True, venkat_K=5.0):
        return make_params(scheme=L.GD_SCHEME_NS, precision=precision, flux=flux, viscous=viscous,
                           venkat_K=venkat_K, venkat_scale=(self.rho_inf, self.u_inf, self.u_inf, self.p_inf),
                           RGAS=self.R, Gamma=self.gamma, mu_ref=self.mu_inf, T_ref=self.T_inf, S_suth=self.S,
                           Pr=self.Pr, bcs=self.bcs())

    def free_stream(self, n):
        return (np.full(n, self.rho_inf), np.full(n, self.u_inf), np.zeros(n), np.full(n, self.p_inf))

    def free_stream_conservative(self, n):
        r, u, v, p = self.free_stream(n)
        U = np.zeros((n, 4), order="F")
        U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v
        U[:, 3] = p / (self.gamma - 1.0) + 0.5 * r * (u * u + v * v)
        return U

    def mach_sensitivity(self, dJ_dU0, dJ_dbc, initial_free_stream=True):
        """dJ/dM_inf at fixed T_inf, p_inf and viscosity (so Re_D moves with
        M_inf): through the inflow state of code -1 and, when the run started
        from the free stream, through the initial state."""
        a_inf = (self.gamma * self.R * self.T_inf) ** 0.5
        d = float(np.dot(dJ_dbc[1], (0.0, a_inf, 0.0, 0.0)))
        if initial_free_stream:
            r, u = self.rho_inf, self.u_inf
            d += float(r * a_inf * dJ_dU0[:, 1].sum() + r * u * a_inf * dJ_dU0[:, 3].sum())
        return d
