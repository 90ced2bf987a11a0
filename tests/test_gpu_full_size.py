"""The 2,079,360-triangle cylinder at full size, through size-independent
properties: the FP64 step is bit-identical whatever the device numbering
(RCM, Hilbert or the seeded random caller order), and the FP32-mixed step stays within
1e-5 of FP64 in density, pressure and wall drag after 100 steps."""
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
