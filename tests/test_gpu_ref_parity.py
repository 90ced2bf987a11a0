"""GPU parity of the reference-scheme operators against the CPU oracle.

Every operator of the reference step runs through the C ABI (via the
``stage`` mirror of the reference's Julia signatures) on the same seeded
inputs as the oracle.  The FP64 kernels follow the reference's operation
order, so the bar is relative L-inf <= 1e-12 per step (measured: bitwise equal
except where a libm function such as pow/sin/exp differs by an ulp), and
<= 1e-9 after 1000 steps of the device-resident loop.
"""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED, conservative, random_fields, rel_linf, tgv_fields

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_1000 = 1e-9


def _thermo(nd=True):
    M = pkg().mesh
    return M.THERMOPHYSICS(RGAS=1.0 if nd else 287.058, Gamma=1.4)


def _tgv_case(n=100):
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(n)
    th = _thermo()
    f = tgv_fields(mesh)
    mv = O.MeshView(mesh)
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    O.cells2nodes(mv, f, U)
    return mesh, mv, th, f, U


def _box_case(quads="mixed", seed=SEED, nx=24, ny=17):
    M = pkg().mesh
    mesh = M.make_box_mesh(nx, ny, 2.0, 1.0, quads=quads)
    th = _thermo()
    f = random_fields(mesh, seed=seed)
    mv = O.MeshView(mesh)
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    O.cells2nodes(mv, f, U)
    return mesh, mv, th, f, U


BOX_BCS = {-1: (O.BC_FIXED, (1.3, 0.7, -0.2, 1.9)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_SLIP_WALL, ()),
           -4: (O.BC_ISOTHERMAL_WALL, (1.1,))}


def _cfg(S, flux=0, limiter=0, bcs=None, energy=O.ENERGY_CONST, mu=1.0 / 1600, k=1.0 / 0.72):
    return S.SchemeConfig(flux=flux, limiter=limiter, bcs=dict(bcs or {}), energy_law=energy, mu_const=mu,
                          k_const=k)


def _oparams(flux=0, limiter=0, bcs=None, energy=O.ENERGY_CONST, mu=1.0 / 1600, k=1.0 / 0.72, udiff=True):
    return O.params(flux=flux, limiter=limiter, include_udiff=udiff, energy_law=energy, mu_const=mu, k_const=k,
                    bcs=bcs)


def test_calc_one_stage_fast_tgv():
    S = pkg().stage
    mesh, mv, th, f, U = _tgv_case()
    Ud, _ = O.calc_diff_term(mv, f, th, _oparams(), U)
    FLo, UNo = O.calc_one_stage_muscl_cpu_fast(mv, f, th, _oparams(), U, Ud, 1.0, 1.3e-3, 0.0)
    ct = pkg().mesh.distibuteCellsInThreadsSA(4, mesh.nCells)
    FL = np.zeros((mesh.nCells, 4), order="F"); UN = np.zeros((mesh.nCells, 4), order="F")
    S.calcOneStageMUSCLcpuFast(ct, 1.0, 1.3e-3, 0.0, mesh, f, th, U, FL, Ud, UN, config=_cfg(S))
    print("FLUXES", rel_linf(FL, FLo), "U", rel_linf(UN, UNo))
    assert rel_linf(FL, FLo) <= TOL_STEP
    assert rel_linf(UN, UNo) <= TOL_STEP


def test_calc_diff_term_tgv_and_air_k():
    S = pkg().stage
    mesh, mv, th, f, U = _tgv_case()
    for energy in (O.ENERGY_CONST, O.ENERGY_AIR_POLY):
        Udo, _ = O.calc_diff_term(mv, f, th, _oparams(energy=energy), U)
        Ud = np.zeros((mesh.nCells, 4), order="F")
        ct = pkg().mesh.distibuteCellsInThreadsSA(2, mesh.nCells)
        S.calcDiffTerm(ct, ct, mesh, f, None, th, U, None, Ud, config=_cfg(S, energy=energy))
        print("Udiff", energy, rel_linf(Ud[:, 1:], Udo[:, 1:]))
        assert rel_linf(Ud[:, 1:], Udo[:, 1:]) <= TOL_STEP
        assert np.all(Ud[:, 0] == 0.0)


@pytest.mark.parametrize("quads", ["none", "mixed", "all"])
@pytest.mark.parametrize("flux", [O.FLUX_AUSM_PLUS, O.FLUX_ROE, O.FLUX_AUSM_PLUS_M2020])
def test_stage_box_boundaries(quads, flux):
    S = pkg().stage
    mesh, mv, th, f, U = _box_case(quads)
    Ud, _ = O.calc_diff_term(mv, f, th, _oparams(bcs=BOX_BCS), U)
    FLo, UNo = O.calc_one_stage_muscl_cpu_fast(mv, f, th, _oparams(flux=flux, bcs=BOX_BCS), U, Ud, 1.0, 1e-3, 0.25)
    ct = pkg().mesh.distibuteCellsInThreadsSA(3, mesh.nCells)
    FL = np.zeros((mesh.nCells, 4), order="F"); UN = np.zeros((mesh.nCells, 4), order="F")
    S.calcOneStageMUSCLcpuFast(ct, 1.0, 1e-3, 0.25, mesh, f, th, U, FL, Ud, UN, config=_cfg(S, flux=flux, bcs=BOX_BCS))
    assert rel_linf(FL, FLo) <= TOL_STEP
    assert rel_linf(UN, UNo) <= TOL_STEP
    # two-phase variant (no Udiff)
    FLo2, UNo2 = O.calc_one_stage_muscl(mv, f, th, _oparams(flux=flux, bcs=BOX_BCS, udiff=False), U, 1.0, 1e-3, 0.25)
    UN2 = np.zeros((mesh.nCells, 4), order="F"); FL2 = np.zeros((mesh.nCells, 4), order="F")
    S.calcOneStageMUSCL(1.0, 1e-3, 0.25, mesh, f, th, ct, U, Ud, UN2, FLUXES=FL2, config=_cfg(S, flux=flux, bcs=BOX_BCS))
    assert rel_linf(UN2, UNo2) <= TOL_STEP
    assert rel_linf(FL2, FLo2) <= TOL_STEP


@pytest.mark.parametrize("limiter", [O.LIM_REF_HARMONIC, O.LIM_VAN_LEER_A, O.LIM_VAN_ALBADA_A, O.LIM_MINMOD])
def test_stage_limiters(limiter):
    S = pkg().stage
    mesh, mv, th, f, U = _box_case("none", seed=SEED + limiter)
    FLo, UNo = O.calc_one_stage_muscl_cpu_fast(mv, f, th, _oparams(limiter=limiter, bcs=BOX_BCS), U, None, 1.0, 1e-3,
                                               0.0)
    ct = pkg().mesh.distibuteCellsInThreadsSA(2, mesh.nCells)
    FL = np.zeros((mesh.nCells, 4), order="F"); UN = np.zeros((mesh.nCells, 4), order="F")
    Z = np.zeros((mesh.nCells, 4), order="F")
    S.calcOneStageMUSCLcpuFast(ct, 1.0, 1e-3, 0.0, mesh, f, th, U, FL, Z, UN,
                               config=_cfg(S, limiter=limiter, bcs=BOX_BCS))
    assert rel_linf(FL, FLo) <= TOL_STEP


def test_tgv_mirror_boundaries():
    """Dimensional TGV on a non-periodic square with the analytic mirror ghost."""
    S = pkg().stage
    M = pkg().mesh
    mesh = M.make_box_mesh(20, 20, 2 * np.pi, 2 * np.pi, quads="none")
    th = M.THERMOPHYSICS()
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
    f.densityCells[:] = 1.0
    f.UxCells[:] = 40 * np.sin(x) * np.cos(y)
    f.UyCells[:] = -40 * np.cos(x) * np.sin(y)
    f.pressureCells[:] = 1e5 + 0.25 * 1600 * (np.cos(2 * x) + np.cos(2 * y))
    mv = O.MeshView(mesh)
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    O.cells2nodes(mv, f, U)
    tgv = {c: (O.BC_TGV_MIRROR, (1e5, 40.0, 1.0, 1e-2)) for c in (-1, -2, -3, -4)}
    FLo, UNo = O.calc_one_stage_muscl_cpu_fast(mv, f, th, _oparams(bcs=tgv, energy=O.ENERGY_AIR_POLY, mu=1e-2), U,
                                               None, 1.0, 1e-5, 0.3)
    ct = M.distibuteCellsInThreadsSA(2, mesh.nCells)
    FL = np.zeros((mesh.nCells, 4), order="F"); UN = np.zeros((mesh.nCells, 4), order="F")
    S.calcOneStageMUSCLcpuFast(ct, 1.0, 1e-5, 0.3, mesh, f, th, U, FL, np.zeros((mesh.nCells, 4), order="F"), UN,
                               config=_cfg(S, bcs=tgv))
    assert rel_linf(FL, FLo) <= 1e-11  # libm sin/cos/exp in the ghost state may differ by an ulp


def test_update_variables_and_cells2nodes():
    S = pkg().stage
    mesh, mv, th, f, U = _box_case("mixed")
    FLo, UNo = O.calc_one_stage_muscl_cpu_fast(mv, f, th, _oparams(bcs=BOX_BCS), U, None, 1.0, 1e-3, 0.0)
    fo = f.copy(); Uo = np.array(U, order="F", copy=True)
    Do = O.update_variables(fo, th.Gamma, UNo, Uo)
    Uno = O.cells2nodes(mv, fo, Uo)
    ct = pkg().mesh.distibuteCellsInThreadsSA(2, mesh.nCells)
    nt = pkg().mesh.distibuteCellsInThreadsSA(2, mesh.nNodes)
    fg = f.copy(); Ug = np.array(U, order="F", copy=True); D = np.zeros_like(Ug)
    S.register_mesh(mesh)
    S.calcOneStageMUSCLcpuFast(ct, 1.0, 1e-3, 0.0, mesh, f, th, U, np.zeros_like(U), np.zeros_like(U),
                               np.zeros_like(U), config=_cfg(S, bcs=BOX_BCS))
    S.updateVariablesSA(ct, th.Gamma, UNo, Ug, D, fg)
    assert rel_linf(D, Do) <= TOL_STEP
    assert np.array_equal(Ug, Uo)
    for name in ("densityCells", "UxCells", "UyCells", "pressureCells", "aSoundCells", "VMAXCells"):
        assert rel_linf(getattr(fg, name), getattr(fo, name)) <= TOL_STEP, name
    Un = np.zeros((mesh.nNodes, 4), order="F")
    S.cells2nodesSolutionReconstructionWithStencils(nt, mesh, fg, None, Ug, Un)
    assert rel_linf(Un, Uno) <= TOL_STEP
    for name in ("densityNodes", "UxNodes", "UyNodes", "pressureNodes"):
        assert rel_linf(getattr(fg, name), getattr(fo, name)) <= TOL_STEP, name


def _oracle_loop(mesh, mv, th, f, U, n_steps, dt, prm_o, viscous=True):
    f = f.copy(); U = np.array(U, order="F", copy=True)
    t = 0.0
    for _ in range(n_steps):
        Ud = O.calc_diff_term(mv, f, th, prm_o, U, n_threads=4)[0] if viscous else None
        _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, prm_o, U, Ud, 1.0, dt, t, n_threads=4)
        O.update_variables(f, th.Gamma, UN, U, n_threads=4)
        O.cells2nodes(mv, f, None, n_threads=4)
        t += dt
    return f, U


def _solver(mesh, th, f, viscous=True, order=None, bcs=None, perm=None):
    S = pkg().stage
    L = pkg()._lib
    p = S.make_params(viscous=viscous, energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72,
                      RGAS=th.RGAS, Gamma=th.Gamma, bcs=bcs or {})
    s = S.GodunovSolver(mesh, p, order=L.GD_ORDER_AS_GIVEN if order is None else order, perm=perm)
    s.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells,
                     nodes=(f.densityNodes, f.UxNodes, f.UyNodes, f.pressureNodes))
    return s


def test_device_loop_matches_oracle_20_steps():
    mesh, mv, th, f, U = _tgv_case(40)
    fo, Uo = _oracle_loop(mesh, mv, th, f, U, 20, 2e-3, _oparams())
    s = _solver(mesh, th, f)
    s.run(20, dt=2e-3)
    assert rel_linf(s.conservative(), Uo) <= TOL_STEP * 20
    rC, uC, vC, pC = s.fields()[:4]
    assert rel_linf(pC, fo.pressureCells) <= TOL_STEP * 20


@pytest.mark.parametrize("order", ["rcm", "hilbert", "random"])
def test_reordering_is_bitwise_invisible(order):
    L = pkg()._lib
    M = pkg().mesh
    mesh, mv, th, f, U = _box_case("mixed", nx=30, ny=20)
    a = _solver(mesh, th, f, bcs=BOX_BCS)
    a.run(15, dt=5e-4)
    if order == "rcm":
        b = _solver(mesh, th, f, bcs=BOX_BCS, order=L.GD_ORDER_RCM)
    elif order == "hilbert":
        b = _solver(mesh, th, f, bcs=BOX_BCS, order=L.GD_ORDER_HILBERT)
    else:
        b = _solver(mesh, th, f, bcs=BOX_BCS, order=L.GD_ORDER_PERMUTATION,
                    perm=M.random_permutation(mesh.nCells, SEED))
    b.run(15, dt=5e-4)
    assert np.array_equal(a.conservative(), b.conservative())


def test_tgv_1000_steps_config0():
    """The 20,000-triangle periodic TGV, 1000 FP64 steps at CFL 0.5 (local
    dt = CFL min h/(|u|+a), h = 2A/perimeter), device-resident loop vs the
    oracle loop.  (The reference's tau, CFL max(HX)/max(VMAX), is 2.4x over
    the stable step on this mesh and blows up.)"""
    L = pkg()._lib
    mesh, mv, th, f, U = _tgv_case(100)
    fo = f.copy()
    fo.aSoundCells[:] = np.sqrt(th.Gamma * fo.pressureCells / fo.densityCells)
    fo.VMAXCells[:] = np.sqrt(fo.UxCells * fo.UxCells + fo.UyCells * fo.UyCells) + fo.aSoundCells
    s = _solver(mesh, th, f)
    n = 1000
    st = s.run(n, cfl=0.5, dt_policy=L.GD_DT_LOCAL)
    t = 0.0
    Uo = np.array(U, order="F", copy=True)
    prm = _oparams()
    for k in range(n):
        dt = O.local_dt(mv, fo, 0.5)
        Ud = O.calc_diff_term(mv, fo, th, prm, Uo, n_threads=4)[0]
        _, UN = O.calc_one_stage_muscl_cpu_fast(mv, fo, th, prm, Uo, Ud, 1.0, dt, t, n_threads=4)
        O.update_variables(fo, th.Gamma, UN, Uo, n_threads=4)
        O.cells2nodes(mv, fo, None, n_threads=4)
        t += dt
    err = rel_linf(s.conservative(), Uo)
    print("1000-step rel Linf", err, "t", t, "last dt", st.last_dt, dt)
    assert err <= TOL_1000
    assert abs(s.flow_time - t) <= 1e-12 * t


@pytest.mark.parametrize("quads", ["none", "mixed", "all"])
def test_two_phase_face_states(quads):
    """Phase one of calcOneStageMUSCL: the mirror fills ULeft1..4 / URight1..4
    as the reference does (computeMUSCLStencilsCUDA.jl:3-28), bit for bit
    with the oracle's face states, slot 4 of triangles untouched."""
    S = pkg().stage
    mesh, mv, th, f, U = _box_case(quads)
    prm = _oparams(bcs=BOX_BCS, udiff=False)
    _, UNo, ULo, URo = O.calc_one_stage_muscl(mv, f, th, prm, U, 1.0, 1e-3, 0.25, face_states=True)
    n = mesh.nCells
    ct = pkg().mesh.distibuteCellsInThreadsSA(3, n)
    UL = [np.full((4, n), 7.0, order="F") for _ in range(4)]
    UR = [np.full((4, n), -7.0, order="F") for _ in range(4)]
    UN = np.zeros((n, 4), order="F")
    S.calcOneStageMUSCL(1.0, 1e-3, 0.25, mesh, f, th, ct, U, None, UN, *UL, *UR, config=_cfg(S, bcs=BOX_BCS))
    assert rel_linf(UN, UNo) <= TOL_STEP
    nv = mesh.mesh_connectivity[:, 2]
    for k in range(4):
        has = nv > k
        assert np.array_equal(UL[k][:, has], ULo[k][:, has])
        assert np.array_equal(UR[k][:, has], URo[k][:, has])
        assert np.all(UL[k][:, ~has] == 7.0) and np.all(UR[k][:, ~has] == -7.0)


def test_linear_field_known_answer_on_device():
    """The device step on a linear field equals the closed form that the
    mathematics fixes (face states q(m) at every face midpoint, FLUXES the
    exact Euler flux summed over the faces), on the interior cells where the
    node means are exact; see test_oracle.linear_field_known_answer."""
    import test_oracle as TO
    S = pkg().stage
    th = _thermo()
    mesh, f, mv, U, sel = TO.linear_field_case(40, 32, 1.0, 0.8)
    idx, qm, Fx = TO.linear_field_known_answer(mesh, sel)
    assert len(idx) > 1000
    n = mesh.nCells
    bcs = {c: (O.BC_EXTRAPOLATE, ()) for c in (-1, -2, -3, -4)}
    ct = pkg().mesh.distibuteCellsInThreadsSA(2, n)
    UL = [np.zeros((4, n), order="F") for _ in range(4)]
    UR = [np.zeros((4, n), order="F") for _ in range(4)]
    UN = np.zeros((n, 4), order="F"); FL = np.zeros((n, 4), order="F")
    dt = 1e-3
    S.calcOneStageMUSCL(1.0, dt, 0.0, mesh, f, th, ct, U, None, UN, *UL, *UR, FLUXES=FL, config=_cfg(S, bcs=bcs))
    for k in range(3):
        assert np.abs(UL[k][:, idx] - qm[k]).max() < 1e-12
        assert np.abs(UR[k][:, idx] - qm[k]).max() < 1e-12
    scale = np.abs(Fx).max()
    assert np.abs(FL[idx] - Fx).max() < 1e-10 * scale
    A = mesh.cell_areas[idx, None]
    assert np.abs(UN[idx] - (U[idx] - dt * Fx / A)).max() < 1e-12 * np.abs(U).max()
    # the oracle agrees bit for bit on the whole mesh
    FLo, _ = O.calc_one_stage_muscl(mv, f, th, _oparams(bcs=bcs, udiff=False), U, 1.0, dt, 0.0)
    assert rel_linf(FL, FLo) <= TOL_STEP


@pytest.mark.parametrize("x32", [False, True])
@pytest.mark.parametrize("quads", ["none", "mixed"])
def test_reference_gpu_stages(quads, x32):
    """calcOneStageCUDA / calcOneStageCUDAx32 through the mirror: given the
    oracle's face states (as computeMUSCLStencilsCUDA hands them over), the
    update and the accumulated iFlux sums equal the oracle's restatement of the
    GPU kernel bit for bit (Float32 arithmetic included); without face states
    the device's own phase one gives the same result."""
    S = pkg().stage
    mesh, mv, th, f, U = _box_case(quads)
    prm = _oparams(bcs=BOX_BCS, udiff=False)
    _, _, ULo, URo = O.calc_one_stage_muscl(mv, f, th, prm, U, 1.0, 1e-3, 0.25, face_states=True)
    Fo, UNo = O.calc_one_stage_cuda(mv, th, 1.0, 1e-3, U, ULo, URo, x32=x32)
    n = mesh.nCells
    ct = pkg().mesh.distibuteCellsInThreadsSA(3, n)
    stage = S.calcOneStageCUDAx32 if x32 else S.calcOneStageCUDA
    UN = np.zeros((n, 4), order="F")
    fl = [np.zeros(n) for _ in range(4)]
    UL = [np.asfortranarray(ULo[k]) for k in range(4)]
    UR = [np.asfortranarray(URo[k]) for k in range(4)]
    stage(1.0, 1e-3, 0.25, mesh, f, th, ct, U, None, UN, *UL, *UR, *fl, config=_cfg(S, bcs=BOX_BCS))
    F = np.stack(fl, 1)
    print("x32" if x32 else "fp64", quads, "iFlux", rel_linf(F, Fo), "U", rel_linf(UN, UNo))
    assert np.array_equal(F, Fo)
    assert np.array_equal(UN, UNo)
    UN2 = np.zeros((n, 4), order="F")
    fl2 = [np.zeros(n) for _ in range(4)]
    stage(1.0, 1e-3, 0.25, mesh, f, th, ct, U, None, UN2, *([None] * 8), *fl2, config=_cfg(S, bcs=BOX_BCS))
    assert np.array_equal(UN2, UN)
    assert np.array_equal(np.stack(fl2, 1), F)


@pytest.mark.parametrize("limiter", [O.LIM_REF_HARMONIC, O.LIM_VAN_ALBADA_A, O.LIM_MINMOD])
@pytest.mark.parametrize("case", ["tgv", "box", "cylinder"])
def test_face_pass_equals_per_side_evaluation(case, limiter, monkeypatch):
    """The reference loop evaluates each interior triangle-triangle face with
    mirrored stencils once and every other face slot (boundary faces, faces of
    quads, faces whose sides do not mirror bit for bit) from its cell's side,
    in the face pass; the step then sums the stored fluxes.  With AUSM+ and an
    odd, symmetric limiter the two sides' evaluations of a mirrored face are
    exact negatives, so 20 steps give the same state, wall force and time as
    evaluating every face from both of its cells in the step
    (GODUNOV_REF_PER_SIDE=1): on a periodic square, on a box with quads,
    walls, inflow and outflow, and on an O-grid cylinder with an oblique slip
    wall."""
    S, L, M = pkg().stage, pkg()._lib, pkg().mesh
    if case == "tgv":
        mesh = M.make_tgv_periodic_mesh(64)
        f = tgv_fields(mesh, noise=1e-3)
        bcs, wall = {}, 0
    elif case == "cylinder":
        mesh = M.make_cylinder_ogrid_mesh(48, 25, r0=0.5, r1=5.0, dr0=0.02)
        f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
        f.densityCells[:] = 1.0 + 1e-3 * M.normal(SEED + 5, mesh.nCells)
        f.UxCells[:], f.UyCells[:], f.pressureCells[:] = 0.3, 0.0, 1.0 / 1.4
        bcs = {-1: (L.GD_BC_FIXED, (1.0, 0.3, 0.0, 1.0 / 1.4)), -2: (L.GD_BC_EXTRAPOLATE, ()),
               -3: (L.GD_BC_SLIP_WALL, ())}
        wall = -3
    else:
        mesh = M.make_box_mesh(40, 24, 2.0, 1.0, quads="mixed")
        f = random_fields(mesh, seed=SEED + 3)
        bcs = {-1: (L.GD_BC_FIXED, (1.3, 0.7, -0.2, 1.9)), -2: (L.GD_BC_EXTRAPOLATE, ()),
               -3: (L.GD_BC_SLIP_WALL, ()), -4: (L.GD_BC_EXTRAPOLATE, ())}
        wall = -3
    prm = S.make_params(limiter=limiter, viscous=True, energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600,
                        k_const=1.0 / 0.72, RGAS=1.0, Gamma=1.4, bcs=bcs)
    out = []
    for per_side in ("0", "1"):
        monkeypatch.setenv("GODUNOV_REF_PER_SIDE", per_side)
        s = S.GodunovSolver(mesh, prm, order=L.GD_ORDER_HILBERT)
        s.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells)
        st = s.run(20, cfl=0.3, dt_policy=L.GD_DT_LOCAL, wall_code=wall)
        out.append((s.conservative(), st.wall_force_sum[0], st.flow_time))
        s.close()
    (Ua, wa, ta), (Ub, wb, tb) = out
    assert np.abs(Ua - f.densityCells[:, None]).max() > 0
    assert np.array_equal(Ua, Ub)
    assert wa == wb and ta == tb
