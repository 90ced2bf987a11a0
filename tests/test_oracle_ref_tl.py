"""The tangent-linear reference loop of the oracle (the reference scheme's
time loop — calcDiffTerm, calcOneStageMUSCLcpuFast, updateVariablesSA, the
node interpolation — instantiated with forward-mode dual numbers), which the
reference-scheme adjoint is checked against.

Pinned by what the mathematics fixes: its value part is the oracle's own
loop of the pointwise functions, bit for bit; its derivative is linear in the
seed (doubling the seed doubles it exactly); and it matches central
differences of the loop, viscous and inviscid, with each limiter — on the
periodic triangle square and on boxes of triangles and quads with every
boundary ghost (fixed state, zero gradient, slip wall, isothermal wall, the
time-dependent TGV mirror), with the summed wall force."""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED, tgv_fields

DT = 2e-3


def _thermo():
    return pkg().mesh.THERMOPHYSICS(RGAS=1.0, Gamma=1.4)


def _case(n=24, noise=1e-2):
    """Periodic TGV with seeded noise on every primitive: the symmetric TGV has
    exactly equal neighbour values along its symmetry lines, where the limiters
    have a kink; there the dual number takes one side's derivative and central
    differences the average of both, so the comparison needs a state without
    such ties."""
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(n)
    th = _thermo()
    f = tgv_fields(mesh, noise=noise)
    r = M.uniform01(SEED + 3, 3 * mesh.nCells).reshape(3, -1) - 0.5
    f.UxCells[:] += noise * r[0]
    f.UyCells[:] += noise * r[1]
    f.pressureCells[:] *= 1.0 + noise * r[2]
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    return mesh, O.MeshView(mesh), th, U


def _oprm(limiter=O.LIM_REF_HARMONIC):
    return O.params(limiter=limiter, include_udiff=True, energy_law=O.ENERGY_CONST, mu_const=1.0 / 1600,
                    k_const=1.0 / 0.72)


def _oracle_loop(mesh, mv, th, U0, n_steps, prm, viscous):
    """The same loop from the oracle's double-precision pointwise functions:
    primitives of U0 and their node means first, then per step the diff term,
    the fast stage, updateVariablesSA and the node interpolation."""
    M = pkg().mesh
    f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
    U = np.array(U0, order="F", copy=True)
    O.update_variables(f, th.Gamma, np.array(U0, order="F", copy=True), U, n_threads=2)
    O.cells2nodes(mv, f, None, n_threads=2)
    for _ in range(n_steps):
        Ud = O.calc_diff_term(mv, f, th, prm, U, n_threads=2)[0] if viscous else None
        _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, prm, U, Ud, 1.0, DT, 0.0, n_threads=2)
        O.update_variables(f, th.Gamma, UN, U, n_threads=2)
        O.cells2nodes(mv, f, None, n_threads=2)
    return U


def _direction(mesh, U, seed):
    M = pkg().mesh
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    r = M.uniform01(seed, 12).reshape(4, 3)
    scale = np.abs(U).max(axis=0)
    d = np.stack([scale[v] * np.sin((1.0 + 2.0 * r[v, 0]) * (np.cos(6.3 * r[v, 1]) * x + np.sin(6.3 * r[v, 1]) * y)
                                    + 6.3 * r[v, 2]) for v in range(4)], 1)
    return np.asfortranarray(d)


@pytest.mark.parametrize("viscous", [True, False])
def test_value_part_is_the_oracle_loop(viscous):
    mesh, mv, th, U0 = _case()
    prm = _oprm()
    UN, _ = O.ref_loop_tl(mv, th, prm, U0, None, DT, 6, viscous=viscous)
    Uo = _oracle_loop(mesh, mv, th, U0, 6, prm, viscous)
    assert np.abs(UN - U0).max() > 0
    assert np.array_equal(UN, Uo)


def test_derivative_is_linear_in_the_seed():
    mesh, mv, th, U0 = _case()
    d = _direction(mesh, U0, SEED + 1)
    _, a = O.ref_loop_tl(mv, th, _oprm(), U0, d, DT, 4)
    _, b = O.ref_loop_tl(mv, th, _oprm(), U0, np.asfortranarray(2.0 * d), DT, 4)
    assert np.abs(a).max() > 0
    assert np.array_equal(b, 2.0 * a)


@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("limiter", [O.LIM_REF_HARMONIC, O.LIM_VAN_LEER_A, O.LIM_VAN_ALBADA_A, O.LIM_MINMOD])
def test_derivative_matches_central_differences(limiter, viscous):
    """<w, dU_N> against (J(U0 + h d) - J(U0 - h d)) / 2h for J = <w, U_N>,
    h = 1e-7 relative to the state's scale (noisy TGV: the limiters' branch
    points are off the sample points)."""
    M = pkg().mesh
    mesh, mv, th, U0 = _case()
    prm = _oprm(limiter)
    n = mesh.nCells
    w = np.asfortranarray(M.uniform01(SEED + 7, 4 * n).reshape(n, 4, order="F") - 0.5)
    errs = []
    for k in range(2):
        d = _direction(mesh, U0, SEED + 11 + k)
        _, dUN = O.ref_loop_tl(mv, th, prm, U0, d, DT, 5, viscous=viscous)
        tl = float((w * dUN).sum())
        h = 1e-7
        Jp = float((w * O.ref_loop_tl(mv, th, prm, np.asfortranarray(U0 + h * d), None, DT, 5, viscous=viscous)[0]).sum())
        Jm = float((w * O.ref_loop_tl(mv, th, prm, np.asfortranarray(U0 - h * d), None, DT, 5, viscous=viscous)[0]).sum())
        fd = (Jp - Jm) / (2 * h)
        scale = float(np.abs(w * dUN).sum())
        errs.append(abs(tl - fd) / scale)
        print(limiter, viscous, k, tl, fd, errs[-1])
    assert max(errs) < 1e-8


def test_unsupported_configurations_are_refused():
    mesh, mv, th, U = _case(8)
    with pytest.raises(RuntimeError):  # Roe
        O.ref_loop_tl(mv, th, O.params(flux=O.FLUX_ROE, energy_law=O.ENERGY_CONST), U, None, DT, 1)
    with pytest.raises(RuntimeError):  # k(T) of air in the viscous term
        O.ref_loop_tl(mv, th, O.params(energy_law=O.ENERGY_AIR_POLY), U, None, DT, 1)


# a box of triangles and quads: fixed inflow (-1), zero gradient (-2), slip
# wall (-3, the bottom) and isothermal wall (-4, the left side)
BOX_BCS = {-1: (O.BC_FIXED, (1.3, 0.7, -0.2, 1.9)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_SLIP_WALL, ()),
           -4: (O.BC_ISOTHERMAL_WALL, (1.1,))}
# a TGV box whose ghosts are the analytic vortex mirrored across the face at
# the step's flow time (data P, U, L, mu)
TGV_BCS = {c: (O.BC_TGV_MIRROR, (71.43, 1.0, 1.0, 0.05)) for c in (-1, -2, -3, -4)}


def _box(quads="mixed", tgv=False):
    from helpers import random_fields
    M = pkg().mesh
    if tgv:
        mesh = M.make_box_mesh(14, 12, 6.283185307179586, 6.283185307179586, quads=quads)
        f = tgv_fields(mesh, noise=1e-2)
        r = M.uniform01(SEED + 5, 3 * mesh.nCells).reshape(3, -1) - 0.5
        f.UxCells[:] += 1e-2 * r[0]
        f.UyCells[:] += 1e-2 * r[1]
        f.pressureCells[:] *= 1.0 + 1e-2 * r[2]
    else:
        mesh = M.make_box_mesh(14, 9, 2.0, 1.0, quads=quads)
        f = random_fields(mesh, seed=SEED + 6)
    th = _thermo()
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    return mesh, O.MeshView(mesh), th, U


def _bprm(limiter=O.LIM_REF_HARMONIC, tgv=False):
    return O.params(limiter=limiter, include_udiff=True, energy_law=O.ENERGY_CONST, mu_const=1.0 / 1600,
                    k_const=1.0 / 0.72, bcs=TGV_BCS if tgv else BOX_BCS)


def _oracle_loop_t(mesh, mv, th, U0, n_steps, prm, viscous, t0):
    """The oracle's loop of pointwise functions with the flow time advancing
    from t0 by DT per step."""
    M = pkg().mesh
    f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
    U = np.array(U0, order="F", copy=True)
    O.update_variables(f, th.Gamma, np.array(U0, order="F", copy=True), U, n_threads=2)
    O.cells2nodes(mv, f, None, n_threads=2)
    t = t0
    for _ in range(n_steps):
        Ud = O.calc_diff_term(mv, f, th, prm, U, n_threads=2)[0] if viscous else None
        _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, prm, U, Ud, 1.0, DT, t, n_threads=2)
        O.update_variables(f, th.Gamma, UN, U, n_threads=2)
        O.cells2nodes(mv, f, None, n_threads=2)
        t = t + DT
    return U


@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("quads,tgv", [("mixed", False), ("none", False), ("all", False), ("mixed", True)])
def test_value_part_on_boxes_is_the_oracle_loop(quads, tgv, viscous):
    mesh, mv, th, U0 = _box(quads, tgv)
    prm = _bprm(tgv=tgv)
    UN, _ = O.ref_loop_tl(mv, th, prm, U0, None, DT, 5, viscous=viscous, flow_time=0.3)
    Uo = _oracle_loop_t(mesh, mv, th, U0, 5, prm, viscous, 0.3)
    assert np.abs(UN - U0).max() > 0
    assert np.array_equal(UN, Uo)


def test_wall_force_is_the_sum_of_the_wall_fluxes():
    """Step 0's wall force equals the sum over the slip wall's faces of the
    oracle's own face fluxes (calcOneStageMUSCL's face states through AUSM+)."""
    mesh, mv, th, U0 = _box("mixed")
    prm = _bprm()
    _, _, F, _ = O.ref_loop_tl_force(mv, th, prm, U0, None, DT, 1, -3, viscous=False)
    M = pkg().mesh
    f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
    U = np.array(U0, order="F", copy=True)
    O.update_variables(f, th.Gamma, np.array(U0, order="F", copy=True), U, n_threads=2)
    O.cells2nodes(mv, f, None, n_threads=2)
    _, _, UL, UR = O.calc_one_stage_muscl(mv, f, th, prm, U, 1.0, DT, 0.0, face_states=True)
    ref = np.zeros(2)
    walls = 0
    for i in range(mesh.nCells):
        for k in range(mesh.mesh_connectivity[i, 2]):
            if mesh.cell_stiffness[i, k] == -3:
                e = O.ausm_plus_flux(UR[k][:, i], UL[k][:, i], mesh.cell_edges_Nx[i, k], mesh.cell_edges_Ny[i, k],
                                     mesh.cell_edges_length[i, k], th.Gamma)
                ref += np.asarray(e)[1:3]
                walls += 1
    assert walls > 0 and np.abs(ref).max() > 0
    assert np.allclose(F, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("quads,tgv,limiter", [("mixed", False, O.LIM_REF_HARMONIC), ("all", False, O.LIM_VAN_ALBADA_A),
                                               ("none", False, O.LIM_MINMOD), ("mixed", True, O.LIM_REF_HARMONIC)])
def test_box_derivative_matches_central_differences(quads, tgv, limiter, viscous):
    """J = <w, U_N> + 0.7 Fx - 0.4 Fy (the summed force on the slip wall, or
    on face code -2 of the TGV box) against central differences at h = 1e-7."""
    M = pkg().mesh
    mesh, mv, th, U0 = _box(quads, tgv)
    prm = _bprm(limiter, tgv)
    n = mesh.nCells
    wc = -2 if tgv else -3
    w = np.asfortranarray(M.uniform01(SEED + 8, 4 * n).reshape(n, 4, order="F") - 0.5)
    fw = np.array([0.7, -0.4])

    def J(U):
        UN, _, F, _ = O.ref_loop_tl_force(mv, th, prm, np.asfortranarray(U), None, DT, 4, wc, viscous, flow_time=0.1)
        return float((w * UN).sum()) + float(fw @ F)
    errs = []
    for k in range(2):
        d = _direction(mesh, U0, SEED + 21 + k)
        _, dUN, _, dF = O.ref_loop_tl_force(mv, th, prm, U0, d, DT, 4, wc, viscous, flow_time=0.1)
        tl = float((w * dUN).sum()) + float(fw @ dF)
        h = 1e-7
        fd = (J(U0 + h * d) - J(U0 - h * d)) / (2 * h)
        scale = float(np.abs(w * dUN).sum()) + float(np.abs(fw * dF).sum())
        errs.append(abs(tl - fd) / scale)
        print(quads, tgv, limiter, viscous, k, tl, fd, errs[-1], dF)
    assert max(errs) < 1e-7


@pytest.mark.parametrize("viscous", [True, False])
def test_fixed_state_derivative_matches_central_differences(viscous):
    """The derivative along the fixed inflow state (code -1's rho, u, v, p)
    against central differences of the loop with that state perturbed, at
    h = 1e-4 (J is about -36 and its derivative 0.013, so smaller steps
    drown in rounding: 4e-11 off at 1e-4, 5e-8 at 1e-7)."""
    M = pkg().mesh
    mesh, mv, th, U0 = _box("mixed")
    n = mesh.nCells
    w = np.asfortranarray(M.uniform01(SEED + 9, 4 * n).reshape(n, 4, order="F") - 0.5)
    fw = np.array([0.3, 0.8])
    base = np.array(BOX_BCS[-1][1])
    db = np.zeros((O.MAX_BC + 1, 4))
    db[1] = (0.4, -0.6, 0.3, 0.9)

    def J(state):
        bcs = dict(BOX_BCS)
        bcs[-1] = (O.BC_FIXED, tuple(state))
        prm = O.params(limiter=O.LIM_REF_HARMONIC, include_udiff=True, energy_law=O.ENERGY_CONST,
                       mu_const=1.0 / 1600, k_const=1.0 / 0.72, bcs=bcs)
        UN, _, F, _ = O.ref_loop_tl_force(mv, th, prm, U0, None, DT, 4, -3, viscous)
        return float((w * UN).sum()) + float(fw @ F)
    _, dUN, _, dF = O.ref_loop_tl_force(mv, th, _bprm(), U0, np.zeros_like(U0), DT, 4, -3, viscous, dbc=db)
    tl = float((w * dUN).sum()) + float(fw @ dF)
    h = 1e-4
    fd = (J(base + h * db[1]) - J(base - h * db[1])) / (2 * h)
    print(viscous, tl, fd, abs(tl - fd) / abs(tl))
    assert abs(tl) > 0
    assert abs(tl - fd) <= 1e-7 * abs(tl)
