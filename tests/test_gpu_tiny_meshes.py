"""The smallest meshes through the same kernels: one quad whose four faces
are all boundary faces (one cell, no interior face), one square split into
two triangles, and a 2 x 1 box with a quad and two triangles.  Every wave of
these launches is mostly lanes past the end of the cells or faces, which the
cooperative gathers must handle (they load stand-in records and store
nothing).  Same bars as the full-size parity tests: FP64 ns step <= 1e-12 of
the oracle, FP32-mixed <= 2e-5 of the float oracle, the adjoint's <dJ/dU0, d>
<= 1e-9 of the dual-number derivative, the reference step <= 1e-12."""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED, random_fields, rel_linf

pytestmark = pytest.mark.gpu

NS_BCS = {-1: (O.BC_FIXED, (1.1, 0.8, 0.1, 1.2)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_ISOTHERMAL_WALL, (1.05,)),
          -4: (O.BC_SLIP_WALL, ())}
GAS = dict(R=1.0, mu_ref=2e-3, T_ref=1.0, S=0.4, Pr=0.72, venkat_K=3.0, venkat_scale=(1.0, 1.0, 1.0, 1.0))
MESHES = [(1, 1, "all"), (1, 1, "none"), (2, 1, "mixed")]
DT = 1e-3


def _mesh(nx, ny, quads):
    return pkg().mesh.make_box_mesh(nx, ny, 1.0, 1.0, quads=quads)


def _state(mesh):
    f = random_fields(mesh, seed=SEED, rho=(0.8, 1.4), vel=(-0.6, 0.9), p=(0.8, 1.6))
    r, u, v, p = f.densityCells, f.UxCells, f.UyCells, f.pressureCells
    U = np.zeros((mesh.nCells, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v; U[:, 3] = p / 0.4 + 0.5 * r * (u * u + v * v)
    return U


def _ns_params(mixed=False, bcs=None):
    S = pkg().stage
    L = pkg()._lib
    return S.make_params(scheme=L.GD_SCHEME_NS, precision=L.GD_FP32_MIXED if mixed else L.GD_FP64,
                         venkat_K=GAS["venkat_K"], venkat_scale=GAS["venkat_scale"], RGAS=GAS["R"], Gamma=1.4,
                         mu_ref=GAS["mu_ref"], T_ref=GAS["T_ref"], S_suth=GAS["S"], Pr=GAS["Pr"],
                         bcs=NS_BCS if bcs is None else bcs)


@pytest.mark.parametrize("nx, ny, quads", MESHES)
def test_tiny_ns_steps_match_oracle(nx, ny, quads):
    S = pkg().stage
    L = pkg()._lib
    mesh = _mesh(nx, ny, quads)
    U = _state(mesh)
    mv = O.MeshView(mesh)
    for mixed in (False, True):
        prm = O.ns_params(gamma=1.4, bcs=NS_BCS, mixed=mixed, **GAS)
        u = U.copy(order="F")
        for _ in range(3):
            u = O.ns_step(mv, prm, DT, u)[0]
        s = S.GodunovSolver(mesh, _ns_params(mixed), order=L.GD_ORDER_HILBERT)
        s.state.set_array(L.GD_ARR_U_OLD, U)
        s.run(3, dt=DT, wall_code=-3)
        got = s.conservative()
        s.close()
        err = rel_linf(got, u)
        print(f"{nx}x{ny} {quads} {'mixed' if mixed else 'fp64'}: {mesh.nCells} cells, rel Linf {err:.2e}")
        assert err <= (2e-5 if mixed else 1e-12)


@pytest.mark.parametrize("objective", ["wall", "terminal"])
@pytest.mark.parametrize("nx, ny, quads", MESHES)
def test_tiny_ns_adjoint_matches_dual_numbers(nx, ny, quads, objective):
    """wall: J = x wall force on code -3 summed over the steps (code -4 made
    zero-gradient: the oracle's force counts every wall face); terminal:
    J = <w, U_N> with the slip wall kept, which checks its transpose."""
    S = pkg().stage
    L = pkg()._lib
    M = pkg().mesh
    mesh = _mesh(nx, ny, quads)
    U = _state(mesh)
    n = mesh.nCells
    bcs = dict(NS_BCS)
    if objective == "wall":
        bcs[-4] = (O.BC_EXTRAPOLATE, ())
        w, fw = None, (1.0, 0.0)
    else:
        w, fw = np.asfortranarray(M.normal(SEED + 3, 4 * n).reshape(n, 4, order="F")), (0.0, 0.0)
    s = S.GodunovSolver(mesh, _ns_params(bcs=bcs))
    s.state.set_array(L.GD_ARR_U_OLD, U)
    g0, gbc, st = s.adjoint(3, DT, checkpoint_every=2, wall_code=-3, force_weight=fw, dJ_dUN=w)
    s.close()
    mv = O.MeshView(mesh)
    prm = O.ns_params(gamma=1.4, bcs=bcs, **GAS)
    d = np.asfortranarray(M.normal(SEED + 7, 4 * n).reshape(n, 4, order="F") * np.abs(U).max(axis=0))
    u, du, dJ = U.copy(order="F"), d.copy(order="F"), 0.0
    for _ in range(3):
        u, du, _, df = O.ns_step_tl(mv, prm, DT, u, du)
        dJ += fw[0] * df[0]
    if w is not None:
        dJ += float((w * du).sum())
    adj = float((g0 * d).sum())
    print(f"{nx}x{ny} {quads} {objective}: <dJ/dU0, d> {adj} dual {dJ}")
    assert abs(adj - dJ) <= 1e-9 * abs(dJ)


@pytest.mark.parametrize("nx, ny, quads", MESHES)
def test_tiny_reference_step_matches_oracle(nx, ny, quads):
    S = pkg().stage
    L = pkg()._lib
    M = pkg().mesh
    mesh = _mesh(nx, ny, quads)
    th = M.THERMOPHYSICS(RGAS=1.0, Gamma=1.4)
    f = random_fields(mesh, seed=SEED)
    mv = O.MeshView(mesh)
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    O.cells2nodes(mv, f, U)
    bcs = {-1: (O.BC_FIXED, (1.3, 0.7, -0.2, 1.9)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_SLIP_WALL, ()),
           -4: (O.BC_EXTRAPOLATE, ())}
    po = O.params(energy_law=O.ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72, bcs=bcs)
    Ud = O.calc_diff_term(mv, f, th, po, U)[0]
    _, UN = O.calc_one_stage_muscl_cpu_fast(mv, f, th, po, U, Ud, 1.0, 1e-3, 0.0)
    p = S.make_params(energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72, RGAS=1.0, bcs=bcs)
    s = S.GodunovSolver(mesh, p, order=L.GD_ORDER_HILBERT)
    s.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells,
                     nodes=(f.densityNodes, f.UxNodes, f.UyNodes, f.pressureNodes))
    s.run(1, dt=1e-3)
    err = rel_linf(s.conservative(), UN)
    s.close()
    print(f"{nx}x{ny} {quads}: reference step rel Linf {err:.2e}")
    assert err <= 1e-12
