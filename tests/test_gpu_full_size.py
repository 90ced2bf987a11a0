"""Full-size workloads through size-independent properties.  The
2,079,360-triangle cylinder: the FP64 step is bit-identical whatever the
device numbering (RCM, Hilbert or the seeded random caller order), and the
FP32-mixed step stays within 1e-5 of FP64 in density, pressure and wall drag
after 100 steps.  The 2,000,000-triangle periodic Taylor-Green square under the
reference scheme: conservation to rounding and numbering independence."""
import numpy as np
import pytest

from conftest import pkg
from helpers import SEED

pytestmark = pytest.mark.gpu

N_STEPS = 100


@pytest.fixture(scope="module")
def cylinder():
    M = pkg().mesh
    mesh = M.make_cylinder_ogrid_mesh(720, 1445, 0.5, 10.0, 2e-4)
    mesh = M.permute_cells(mesh, M.random_permutation(mesh.nCells, SEED))
    assert mesh.nCells == 2079360
    return mesh


def _run(mesh, precision, order):
    S = pkg().stage
    L = pkg()._lib
    case = S.CylinderCase()
    s = S.GodunovSolver(mesh, case.params(precision=precision), order=order)
    s.state.set_array(L.GD_ARR_U_OLD, case.free_stream_conservative(mesh.nCells))
    st = s.run(N_STEPS, cfl=0.4, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    f = s.fields()
    U = s.conservative()
    s.close()
    return U, f[0], f[3], st.wall_force_sum[0]


def test_full_size_numbering_and_mixed_precision(cylinder):
    L = pkg()._lib
    U_rcm, r64, p64, d64 = _run(cylinder, L.GD_FP64, L.GD_ORDER_RCM)
    U_giv, _, _, d_giv = _run(cylinder, L.GD_FP64, L.GD_ORDER_AS_GIVEN)
    U_hil, _, _, d_hil = _run(cylinder, L.GD_FP64, L.GD_ORDER_HILBERT)
    assert np.array_equal(U_rcm, U_giv)
    assert np.array_equal(U_rcm, U_hil)
    # the wall sum's order follows the layout
    assert abs(d64 - d_giv) <= 1e-12 * abs(d64) and abs(d64 - d_hil) <= 1e-12 * abs(d64)
    _, r32, p32, d32 = _run(cylinder, L.GD_FP32_MIXED, L.GD_ORDER_HILBERT)
    er = np.abs(r32 - r64).max() / np.abs(r64).max()
    ep = np.abs(p32 - p64).max() / np.abs(p64).max()
    ed = abs(d32 - d64) / abs(d64)
    print(f"full size, {N_STEPS} steps: mixed vs fp64 rho {er:.2e} p {ep:.2e} drag {ed:.2e}")
    assert er <= 1e-5 and ep <= 1e-5 and ed <= 1e-5


N_REF = 50


@pytest.fixture(scope="module")
def tgv():
    mesh = pkg().mesh.make_tgv_periodic_mesh(1000)
    assert mesh.nCells == 2000000
    return mesh


def _run_ref(mesh, order):
    M, S, L = pkg().mesh, pkg().stage, pkg()._lib
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    rho = 1.0 + 1e-6 * M.normal(SEED, mesh.nCells)
    u, v = np.sin(x) * np.cos(y), -np.cos(x) * np.sin(y)
    p = 1.0 / (1.4 * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) / 4.0
    prm = S.make_params(viscous=True, energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72,
                        RGAS=1.0, Gamma=1.4, bcs={})
    s = S.GodunovSolver(mesh, prm, order=order)
    s.set_primitives(rho, u, v, p)
    U0 = s.conservative()
    s.run(N_REF, cfl=0.5, dt_policy=L.GD_DT_LOCAL)
    U = s.conservative()
    s.close()
    return U0, U


def test_full_size_reference_scheme_conservation_and_numbering(tgv):
    """The reference scheme (MUSCL, Minmod_Limiter, AUSM+, the calcDiffTerm
    term with constant k) on the benchmark's 2M-triangle periodic TGV: after 50
    steps the totals of rho, rho u, rho v and E are unchanged to rounding (every
    face is evaluated from both sides with mirrored stencils on this
    all-triangle mesh, and the Green-Gauss sums of the Laplacian telescope), and
    the state is bit-identical with the device layout in Hilbert order or as
    given."""
    L = pkg()._lib
    U0, Uh = _run_ref(tgv, L.GD_ORDER_HILBERT)
    _, Ug = _run_ref(tgv, L.GD_ORDER_AS_GIVEN)
    assert np.array_equal(Uh, Ug)
    assert np.abs(Uh - U0).max() > 1e-6
    A = np.asarray(tgv.cell_areas)
    for q in range(4):
        drift = abs((A * Uh[:, q]).sum() - (A * U0[:, q]).sum())
        scale = (A * np.abs(U0[:, q])).sum()
        print(f"reference scheme, {N_REF} steps: total {q} drifts {drift / scale:.2e}")
        assert drift <= 1e-12 * scale
