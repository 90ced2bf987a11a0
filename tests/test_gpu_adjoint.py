"""Discrete adjoint of the ns loop (gd_adjoint_run) on the GPU.

The adjoint is the transpose of the tangent-linear step, so for any direction
d the dot product <dJ/dU0, d> must equal the derivative of J along d that the
oracle's dual-number step computes (FP64: to 1e-9 relative), and the same for
the fixed inflow state.  On the supersonic cylinder the Mach-number derivative
and random directions are also checked against central differences of the
GPU forward loop.  J = sum over steps of the x wall force on code -3 plus
<w, U_N>.
"""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED

pytestmark = pytest.mark.gpu

# only code -3 is a wall here, so the oracle's wall force sums the same faces
BCS = {-1: (O.BC_FIXED, (1.1, 0.8, 0.1, 1.2)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_ISOTHERMAL_WALL, (1.05,)),
       -4: (O.BC_EXTRAPOLATE, ())}
GAS = dict(R=1.0, mu_ref=2e-3, T_ref=1.0, S=0.4, Pr=0.72, venkat_K=3.0, venkat_scale=(1.0, 1.0, 1.0, 1.0))
DT = 1e-3


def _box(quads="mixed"):
    M = pkg().mesh
    mesh = M.make_box_mesh(16, 10, 2.0, 1.0, quads=quads)
    n = mesh.nCells
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    u = M.uniform01(SEED, 4 * n).reshape(4, n)
    r = 1.0 + 0.2 * np.sin(2 * x) * np.cos(3 * y) + 0.02 * u[0]
    vx = 0.6 + 0.3 * np.cos(x + y) + 0.02 * u[1]
    vy = 0.1 * np.sin(3 * x) + 0.02 * u[2]
    p = 1.2 + 0.2 * np.cos(2 * x - y) + 0.02 * u[3]
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * vx; U[:, 2] = r * vy; U[:, 3] = p / 0.4 + 0.5 * r * (vx * vx + vy * vy)
    return mesh, U


def _dev_params(mixed=False, viscous=True, roe=False):
    S = pkg().stage
    L = pkg()._lib
    return S.make_params(scheme=L.GD_SCHEME_NS, precision=L.GD_FP32_MIXED if mixed else L.GD_FP64,
                         flux=L.GD_FLUX_ROE if roe else L.GD_FLUX_AUSM_PLUS, viscous=viscous, venkat_K=GAS["venkat_K"], venkat_scale=GAS["venkat_scale"],
                         RGAS=GAS["R"], Gamma=1.4, mu_ref=GAS["mu_ref"], T_ref=GAS["T_ref"], S_suth=GAS["S"],
                         Pr=GAS["Pr"], bcs=BCS)


def _tl_objective(mesh, U, n, dU, dbc, w, viscous=True, roe=False):
    """Oracle dual-number run: J and dJ along (dU, dbc)."""
    mv = O.MeshView(mesh)
    prm = O.ns_params(gamma=1.4, viscous=viscous, bcs=BCS, flux=O.FLUX_ROE if roe else O.FLUX_AUSM_PLUS, **GAS)
    u, du = U.copy(order="F"), dU.copy(order="F")
    J = dJ = 0.0
    for _ in range(n):
        u, du, f, df = O.ns_step_tl(mv, prm, DT, u, du, dbc)
        J += f[0]
        dJ += df[0]
    return J + float((w * u).sum()), dJ + float((w * du).sum())


def _adjoint(mesh, U, n, c, w, mixed=False, viscous=True, order=None, perm=None, budget=0, roe=False):
    S = pkg().stage
    L = pkg()._lib
    s = S.GodunovSolver(mesh, _dev_params(mixed, viscous, roe), order=L.GD_ORDER_AS_GIVEN if order is None else order,
                        perm=perm)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    g0, gbc, st = s.adjoint(n, DT, checkpoint_every=c, wall_code=-3, force_weight=(1.0, 0.0), dJ_dUN=w,
                            record_budget=budget)
    UN = s.conservative()
    s.close()
    return g0, gbc, st, UN


@pytest.mark.parametrize("flux", ["ausm", "roe"])
@pytest.mark.parametrize("quads", ["none", "mixed"])
@pytest.mark.parametrize("viscous", [False, True])
def test_adjoint_fp64_dot_product_with_tangent_linear_oracle(quads, viscous, flux):
    roe = flux == "roe"
    mesh, U = _box(quads)
    n = mesh.nCells
    M = pkg().mesh
    w = np.asfortranarray(1e-3 * M.normal(SEED + 1, 4 * n).reshape(n, 4, order="F"))
    g0, gbc, st, UN = _adjoint(mesh, U, 6, 4, w, viscous=viscous, roe=roe)
    print("adjoint", flux, quads, viscous, "J", st.J, "recomputed", st.recomputed_steps)
    assert st.recomputed_steps == 4
    for k in range(3):
        d = np.asfortranarray(M.normal(SEED + 10 + k, 4 * n).reshape(n, 4, order="F") * np.abs(U).max(axis=0))
        J, dJ = _tl_objective(mesh, U, 6, d, None, w, viscous, roe)
        adj = float((g0 * d).sum())
        print("  <dJ/dU0, d>", adj, "tangent", dJ)
        assert abs(adj - dJ) <= 1e-9 * max(abs(dJ), 1e-300)
    assert abs(st.J - J) <= 1e-11 * abs(J)
    # the fixed inflow state (code -1)
    dbc = np.zeros((O.MAX_BC + 1, 4))
    dbc[1] = (0.3, -0.5, 0.7, 0.2)
    _, dJb = _tl_objective(mesh, U, 6, np.zeros_like(U), dbc, w, viscous, roe)
    assert abs(float((gbc * dbc).sum()) - dJb) <= 1e-9 * abs(dJb)


def test_adjoint_periodic_tgv_mesh_dot_product():
    """The periodic 20,000-triangle Taylor-Green mesh (no boundary faces):
    J = <w, U_N> after 8 viscous FP64 steps, against the dual-number oracle."""
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(100)
    n = mesh.nCells
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    r = 1.0 + 1e-3 * M.normal(SEED, n)
    u, v = np.sin(x) * np.cos(y), -np.cos(x) * np.sin(y)
    p = 1.0 / (1.4 * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) / 4.0
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v; U[:, 3] = p / 0.4 + 0.5 * r * (u * u + v * v)
    w = np.asfortranarray(M.normal(SEED + 3, 4 * n).reshape(n, 4, order="F"))
    g0, _, st, _ = _adjoint(mesh, U, 8, 3, w)
    d = np.asfortranarray(M.normal(SEED + 4, 4 * n).reshape(n, 4, order="F") * np.abs(U).max(axis=0))
    J, dJ = _tl_objective(mesh, U, 8, d, None, w)
    adj = float((g0 * d).sum())
    print("tgv <dJ/dU0, d>", adj, "tangent", dJ)
    assert abs(adj - dJ) <= 1e-9 * abs(dJ)
    assert abs(st.J - J) <= 1e-11 * abs(J)


def test_adjoint_independent_of_checkpointing_and_numbering():
    L = pkg()._lib
    M = pkg().mesh
    mesh, U = _box()
    a = _adjoint(mesh, U, 5, 5, None)
    b = _adjoint(mesh, U, 5, 1, None)
    c = _adjoint(mesh, U, 5, 2, None, order=L.GD_ORDER_PERMUTATION, perm=M.random_permutation(mesh.nCells, SEED))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # records of every segment kept (no recomputation), or of two of the five
    d = _adjoint(mesh, U, 5, 1, None, budget=1 << 40)
    e = _adjoint(mesh, U, 5, 1, None, budget=2 * 16 * 8 * mesh.nCells)
    assert d[2].recomputed_steps == 0 and e[2].recomputed_steps == 3
    assert np.array_equal(a[0], d[0]) and np.array_equal(a[0], e[0]) and np.array_equal(a[1], e[1])
    # another device numbering sums the same terms in another order
    assert np.allclose(a[0], c[0], rtol=1e-12, atol=1e-13 * np.abs(a[0]).max())
    assert np.array_equal(a[3], b[3])


@pytest.mark.parametrize("flux", ["ausm", "roe"])
def test_adjoint_mixed_close_to_fp64(flux):
    mesh, U = _box()
    n = mesh.nCells
    g64 = _adjoint(mesh, U, 6, 3, None, roe=flux == "roe")[0]
    g32, _, st, _ = _adjoint(mesh, U, 6, 3, None, mixed=True, roe=flux == "roe")
    M = pkg().mesh
    d = np.asfortranarray(M.normal(SEED + 5, 4 * n).reshape(n, 4, order="F") * np.abs(U).max(axis=0))
    a, b = float((g64 * d).sum()), float((g32 * d).sum())
    print("mixed vs fp64 directional derivative", a, b)
    assert abs(a - b) <= 1e-4 * abs(a)


def test_adjoint_ref_scheme_with_air_conductivity_is_refused():
    """The reference scheme's adjoint covers constant mu and k; the viscous
    term with k(T) of air is refused, not approximated."""
    S = pkg().stage
    L = pkg()._lib
    mesh, U = _box()
    s = S.GodunovSolver(mesh, S.make_params(bcs={}))  # the reference scheme, k(T) of air
    s.state.set_array(L.GD_ARR_U_OLD, U)
    with pytest.raises(L.GodunovError, match="reference-scheme adjoint covers"):
        s.adjoint(2, DT)
    s.close()


def noisy_free_stream(case, n, seed, rel=1e-6):
    """The case's free stream with the density scaled by 1 + rel N(0, 1) per
    cell (velocity and pressure unchanged), so that no two cells are equal."""
    M = pkg().mesh
    r, u, v, p = case.free_stream(n)
    r = r * (1.0 + rel * M.normal(seed, n))
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v
    U[:, 3] = p / (case.gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return U


def _cylinder(mach=None):
    S = pkg().stage
    case = S.CylinderCase()
    if mach is not None:
        case.mach = mach
    return case


def _cyl_forward(mesh, case, base, U0, n, dt):
    """Sum over n fixed-dt steps of the x wall force, inflow of ``case`` but
    viscosity and limiter scales of ``base``."""
    S = pkg().stage
    L = pkg()._lib
    prm = base.params()
    bc = prm.bc[1]
    bc.data[0], bc.data[1], bc.data[2], bc.data[3] = case.rho_inf, case.u_inf, 0.0, case.p_inf
    s = S.GodunovSolver(mesh, prm)
    s.state.set_array(L.GD_ARR_U_OLD, U0)
    st = s.run(n, dt=dt, wall_code=-3)
    s.close()
    return st.wall_force_sum[0]


def test_adjoint_cylinder_against_central_differences():
    """Supersonic cylinder on a coarse O-grid, 20 FP64 steps from the free
    stream, checkpoints every 5: dJ/dM_inf (through the inflow and the
    initial state, at fixed viscosity) against central differences of the
    forward loop; random directions against the oracle's dual-number run; and
    16 random directions of a developed state (20 steps from a free stream
    with 1e-6 density noise), for the next 20 steps, against central
    differences (1e-4 relative)."""
    S = pkg().stage
    L = pkg()._lib
    M = pkg().mesh
    base = _cylinder()
    mesh = M.make_cylinder_ogrid_mesh(72, 41, 0.5, 10.0, 2e-3)
    n = mesh.nCells
    U0 = base.free_stream_conservative(n)
    mv = O.MeshView(mesh)
    bcs = {-1: (O.BC_FIXED, (base.rho_inf, base.u_inf, 0.0, base.p_inf)), -2: (O.BC_EXTRAPOLATE, ()),
           -3: (O.BC_ISOTHERMAL_WALL, (base.T_wall,))}
    prm = O.ns_params(gamma=1.4, R=base.R, Pr=base.Pr, mu_ref=base.mu_inf, T_ref=base.T_inf, S=base.S,
                      venkat_K=5.0, venkat_scale=(base.rho_inf, base.u_inf, base.u_inf, base.p_inf), bcs=bcs)
    dt = O.ns_dt(mv, prm, U0, 0.4)
    N = 20
    s = S.GodunovSolver(mesh, base.params())
    s.state.set_array(L.GD_ARR_U_OLD, U0)
    g0, gbc, st = s.adjoint(N, dt, checkpoint_every=5, wall_code=-3)
    s.close()
    J0 = _cyl_forward(mesh, base, base, U0, N, dt)
    assert abs(st.J - J0) <= 1e-12 * abs(J0)
    dJdM = base.mach_sensitivity(g0, gbc)
    h = 1e-6 * base.mach
    Jp = _cyl_forward(mesh, _cylinder(base.mach + h), base, _cylinder(base.mach + h).free_stream_conservative(n), N, dt)
    Jm = _cyl_forward(mesh, _cylinder(base.mach - h), base, _cylinder(base.mach - h).free_stream_conservative(n), N, dt)
    fd = (Jp - Jm) / (2 * h)
    print("dJ/dM adjoint", dJdM, "central difference", fd)
    assert abs(dJdM - fd) <= 1e-5 * abs(fd)
    # the same adjoint against the oracle's dual-number run in random
    # directions (the tangent takes the same side of every kink)
    for k in range(2):
        d = np.asfortranarray(M.normal(SEED + 100 + k, 4 * n).reshape(n, 4, order="F") * np.abs(U0).max(axis=0))
        u, du, dJ = U0.copy(order="F"), d.copy(order="F"), 0.0
        for _ in range(N):
            u, du, f, df = O.ns_step_tl(mv, prm, dt, u, du)
            dJ += df[0]
        adj = float((g0 * d).sum())
        print("direction", k, "adjoint", adj, "tangent-linear oracle", dJ)
        assert abs(adj - dJ) <= 1e-9 * abs(dJ)
    # central differences in random directions need a point where J is
    # differentiable along them.  A developed state grown from the exact free
    # stream borders cells that are still exactly uniform, where the
    # limiter's min/max over equal neighbours and AUSM+'s min(a_L, a_R) of
    # equal states are ties: J is only one-sidedly differentiable there (the
    # adjoint takes the tangent's side, pinned above).  So the developed state
    # grows from the free stream with seeded 1e-6 relative density noise,
    # which breaks every tie, and the directions cover every cell.  The
    # limiter is strongly curved where neighbour differences are small, so
    # the step is 1e-10: the oracle's own central differences are 5 % off at
    # 1e-7, 2e-4 at 1e-9, 1e-5 at 1e-10 (round-off takes over at 1e-11).
    # 16 directions.
    Un = noisy_free_stream(base, n, SEED + 7)
    s = S.GodunovSolver(mesh, base.params())
    s.state.set_array(L.GD_ARR_U_OLD, Un)
    s.run(N, dt=dt)
    U1 = s.conservative()
    g1, _, st = s.adjoint(N, dt, checkpoint_every=5, wall_code=-3)
    s.close()
    J1 = _cyl_forward(mesh, base, base, U1, N, dt)
    assert abs(st.J - J1) <= 1e-12 * abs(J1)
    scale = np.abs(U1).max(axis=0)
    errs = []
    for k in range(16):
        d = np.asfortranarray(M.normal(SEED + 200 + k, 4 * n).reshape(n, 4, order="F") * scale)
        eps = 1e-10
        Jp = _cyl_forward(mesh, base, base, np.asfortranarray(U1 + eps * d), N, dt)
        Jm = _cyl_forward(mesh, base, base, np.asfortranarray(U1 - eps * d), N, dt)
        fd = (Jp - Jm) / (2 * eps)
        adj = float((g1 * d).sum())
        print("direction", k, "adjoint", adj, "central difference", fd)
        errs.append(abs(adj - fd) / max(abs(fd), 1e-8 * abs(J1)))
    print("max relative difference", max(errs))
    assert max(errs) <= 1e-4


def test_adjoint_kept_on_device():
    """dJ/dU0 can stay in HBM: the run without a host array gives the same
    objective, and gd_state_device_adjoint hands out its device pointer (an
    error before any adjoint has run)."""
    import ctypes as C
    S = pkg().stage
    L = pkg()._lib
    mesh, U = _box()
    s = S.GodunovSolver(mesh, _dev_params())
    ptr = C.POINTER(C.c_double)()
    assert L.load().gd_state_device_adjoint(s.state.handle, C.byref(ptr)) == L.GD_ERR_INVALID
    s.state.set_array(L.GD_ARR_U_OLD, U)
    g_dev, _, st_dev = s.adjoint(4, DT, checkpoint_every=2, wall_code=-3, to_host=False)
    assert g_dev is None
    assert L.load().gd_state_device_adjoint(s.state.handle, C.byref(ptr)) == 0 and bool(ptr)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    g, _, st = s.adjoint(4, DT, checkpoint_every=2, wall_code=-3)
    s.close()
    assert st.J == st_dev.J and np.abs(g).max() > 0
