"""GPU parity of the ns scheme (cell gradients + Venkatakrishnan + AUSM+/Roe +
Sutherland viscous fluxes + isothermal wall) against the CPU oracle.

FP64: the device kernels follow the oracle's operation order without FMA
contraction, so the bar is relative L-inf <= 1e-12 per step (only pow() in the
Sutherland law may differ by an ulp).  FP32-mixed: against the float oracle
<= 2e-5 per step and against the FP64 oracle <= 5e-4 per step (single-precision
slopes and fluxes, FP64 accumulation and update).
"""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED, rel_linf, random_fields

pytestmark = pytest.mark.gpu


def _box(quads="mixed", nx=24, ny=16):
    M = pkg().mesh
    return M.make_box_mesh(nx, ny, 2.0, 1.0, quads=quads)


def _box_state(mesh, g=1.4):
    f = random_fields(mesh, seed=SEED, rho=(0.8, 1.4), vel=(-0.6, 0.9), p=(0.8, 1.6))
    U = np.zeros((mesh.nCells, 4), order="F")
    r, u, v, p = f.densityCells, f.UxCells, f.UyCells, f.pressureCells
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v; U[:, 3] = p / (g - 1) + 0.5 * r * (u * u + v * v)
    return U


BCS = {-1: (O.BC_FIXED, (1.1, 0.8, 0.1, 1.2)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_ISOTHERMAL_WALL, (1.05,)),
       -4: (O.BC_SLIP_WALL, ())}
GAS = dict(R=1.0, mu_ref=2e-3, T_ref=1.0, S=0.4, Pr=0.72, venkat_K=3.0, venkat_scale=(1.0, 1.0, 1.0, 1.0))


def _dev_params(flux=0, viscous=True, mixed=False):
    S = pkg().stage
    L = pkg()._lib
    return S.make_params(scheme=L.GD_SCHEME_NS, precision=L.GD_FP32_MIXED if mixed else L.GD_FP64, flux=flux,
                         viscous=viscous, venkat_K=GAS["venkat_K"], venkat_scale=GAS["venkat_scale"], RGAS=GAS["R"],
                         Gamma=1.4, mu_ref=GAS["mu_ref"], T_ref=GAS["T_ref"], S_suth=GAS["S"], Pr=GAS["Pr"], bcs=BCS)


def _orc_params(flux=0, viscous=True, mixed=False):
    return O.ns_params(gamma=1.4, flux=flux, viscous=viscous, mixed=mixed, bcs=BCS, **GAS)


def _device_steps(mesh, U, n, dt, params, order=None, perm=None, wall_code=0):
    S = pkg().stage
    L = pkg()._lib
    s = S.GodunovSolver(mesh, params, order=L.GD_ORDER_AS_GIVEN if order is None else order, perm=perm)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    st = s.run(n, dt=dt, wall_code=wall_code)
    out = s.conservative()
    s.close()
    return out, st


@pytest.mark.parametrize("quads", ["none", "mixed"])
@pytest.mark.parametrize("flux", [O.FLUX_AUSM_PLUS, O.FLUX_ROE])
@pytest.mark.parametrize("viscous", [False, True])
def test_ns_fp64_one_step(quads, flux, viscous):
    mesh = _box(quads)
    U = _box_state(mesh)
    mv = O.MeshView(mesh)
    UNo, _, _, fo = O.ns_step(mv, _orc_params(flux, viscous), 2e-3, U)
    UN, st = _device_steps(mesh, U, 1, 2e-3, _dev_params(flux, viscous), wall_code=-3)
    err = rel_linf(UN, UNo)
    print("ns fp64 one step", quads, flux, viscous, err)
    assert err <= 1e-12


def test_ns_fp64_ten_steps_and_wall_force():
    """Ten FP64 steps, and the wall force of the last step on the isothermal
    wall (code -3) against the oracle's sum over the same faces (the slip
    walls of the box are made extrapolating so that only -3 is a wall)."""
    bcs = {**BCS, -4: (O.BC_EXTRAPOLATE, ())}
    S = pkg().stage
    L = pkg()._lib
    mesh = _box("mixed")
    U = _box_state(mesh)
    mv = O.MeshView(mesh)
    prm = O.ns_params(gamma=1.4, bcs=bcs, **GAS)
    Uo = U.copy(order="F")
    for _ in range(10):
        Uo, _, _, fo = O.ns_step(mv, prm, 1e-3, Uo)
    dp = S.make_params(scheme=L.GD_SCHEME_NS, precision=L.GD_FP64, viscous=True, venkat_K=GAS["venkat_K"],
                       venkat_scale=GAS["venkat_scale"], RGAS=GAS["R"], Gamma=1.4, mu_ref=GAS["mu_ref"],
                       T_ref=GAS["T_ref"], S_suth=GAS["S"], Pr=GAS["Pr"], bcs=bcs)
    UN, st = _device_steps(mesh, U, 10, 1e-3, dp, wall_code=-3)
    assert rel_linf(UN, Uo) <= 1e-11
    # forces from the same state, summed in a different order
    assert np.allclose([st.wall_force[0], st.wall_force[1]], fo, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("flux", [O.FLUX_AUSM_PLUS, O.FLUX_ROE])
def test_ns_mixed_vs_float_and_fp64_oracle(flux):
    mesh = _box("mixed")
    U = _box_state(mesh)
    mv = O.MeshView(mesh)
    UNf, _, _, _ = O.ns_step(mv, _orc_params(flux, True, mixed=True), 2e-3, U)
    UNd, _, _, _ = O.ns_step(mv, _orc_params(flux, True, mixed=False), 2e-3, U)
    UN, _ = _device_steps(mesh, U, 1, 2e-3, _dev_params(flux, True, mixed=True))
    # compare the increments, the part that carries the single-precision error
    e_f = rel_linf(UN - U, UNf - U)
    e_d = rel_linf(UN - U, UNd - U)
    print("mixed vs float oracle", e_f, "vs fp64 oracle", e_d)
    assert e_f <= 2e-5
    assert e_d <= 5e-4


def test_ns_reordering_invisible_fp64():
    L = pkg()._lib
    M = pkg().mesh
    mesh = _box("mixed", 30, 20)
    U = _box_state(mesh)
    a, _ = _device_steps(mesh, U, 8, 1e-3, _dev_params())
    b, _ = _device_steps(mesh, U, 8, 1e-3, _dev_params(), order=L.GD_ORDER_RCM)
    c, _ = _device_steps(mesh, U, 8, 1e-3, _dev_params(), order=L.GD_ORDER_PERMUTATION,
                         perm=M.random_permutation(mesh.nCells, SEED))
    d, _ = _device_steps(mesh, U, 8, 1e-3, _dev_params(), order=L.GD_ORDER_HILBERT)
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)
    assert np.array_equal(a, d)


def test_ns_free_stream_preserved_periodic():
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(24)
    n = mesh.nCells
    U = np.zeros((n, 4), order="F")
    U[:, 0] = 1.2; U[:, 1] = 1.2 * 0.3; U[:, 2] = -1.2 * 0.2; U[:, 3] = 2.0 / 0.4 + 0.5 * 1.2 * 0.13
    UN, _ = _device_steps(mesh, U, 5, 1e-3, _dev_params())
    assert rel_linf(UN, U) <= 1e-14
    UN, _ = _device_steps(mesh, U, 5, 1e-3, _dev_params(mixed=True))
    assert rel_linf(UN, U) <= 1e-7


def test_ns_small_cylinder_fp64_vs_oracle_50_steps():
    """The supersonic cylinder set-up on a coarse O-grid: 50 FP64 steps with
    the CFL-0.4 viscous dt, device vs oracle (same dt each step)."""
    M = pkg().mesh
    S = pkg().stage
    L = pkg()._lib
    case = S.CylinderCase()
    mesh = M.make_cylinder_ogrid_mesh(72, 41, 0.5, 10.0, 2e-3)
    mv = O.MeshView(mesh)
    n = mesh.nCells
    r, u, v, p = case.free_stream(n)
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v; U[:, 3] = p / 0.4 + 0.5 * r * (u * u + v * v)
    bcs = {-1: (O.BC_FIXED, (case.rho_inf, case.u_inf, 0.0, case.p_inf)), -2: (O.BC_EXTRAPOLATE, ()),
           -3: (O.BC_ISOTHERMAL_WALL, (case.T_wall,))}
    prm = O.ns_params(gamma=1.4, R=case.R, Pr=case.Pr, mu_ref=case.mu_inf, T_ref=case.T_inf, S=case.S,
                      venkat_K=5.0, venkat_scale=(case.rho_inf, case.u_inf, case.u_inf, case.p_inf), bcs=bcs)
    s = S.GodunovSolver(mesh, case.params())
    s.state.set_array(L.GD_ARR_U_OLD, U)
    Uo = U.copy(order="F")
    t = 0.0
    for k in range(50):
        dt = O.ns_dt(mv, prm, Uo, 0.4)
        st = s.run(1, cfl=0.4, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
        assert abs(st.last_dt - dt) <= 1e-12 * dt, (k, st.last_dt, dt)
        Uo, _, _, f = O.ns_step(mv, prm, dt, Uo)
        t += dt
    err = rel_linf(s.conservative(), Uo)
    print("cylinder 50 steps fp64", err, "drag", st.wall_force[0], f[0])
    assert err <= 1e-11
    assert abs(st.wall_force[0] - f[0]) <= 1e-9 * max(abs(f[0]), case.p_inf)


def test_ns_small_cylinder_fp64_1000_steps():
    """1000 FP64 steps of the cylinder set-up (coarse O-grid, CFL-0.4 viscous
    dt chosen on the device each step) against the oracle stepping with its
    own dt: relative L-inf <= 1e-9 after 1000 steps, and the time reached."""
    M = pkg().mesh
    S = pkg().stage
    L = pkg()._lib
    case = S.CylinderCase()
    mesh = M.make_cylinder_ogrid_mesh(48, 25, 0.5, 10.0, 5e-3)
    mv = O.MeshView(mesh)
    n = mesh.nCells
    U = case.free_stream_conservative(n)
    bcs = {-1: (O.BC_FIXED, (case.rho_inf, case.u_inf, 0.0, case.p_inf)), -2: (O.BC_EXTRAPOLATE, ()),
           -3: (O.BC_ISOTHERMAL_WALL, (case.T_wall,))}
    prm = O.ns_params(gamma=1.4, R=case.R, Pr=case.Pr, mu_ref=case.mu_inf, T_ref=case.T_inf, S=case.S,
                      venkat_K=5.0, venkat_scale=(case.rho_inf, case.u_inf, case.u_inf, case.p_inf), bcs=bcs)
    s = S.GodunovSolver(mesh, case.params(), order=L.GD_ORDER_RCM)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    st = s.run(1000, cfl=0.4, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    UN = s.conservative()
    s.close()
    O.set_threads(8)
    Uo, t = U.copy(order="F"), 0.0
    for _ in range(1000):
        dt = O.ns_dt(mv, prm, Uo, 0.4)
        Uo, _, _, _ = O.ns_step(mv, prm, dt, Uo)
        t += dt
    err = rel_linf(UN, Uo)
    print("cylinder 1000 steps fp64", err, "time", st.flow_time, t)
    assert err <= 1e-9
    assert abs(st.flow_time - t) <= 1e-12 * t


@pytest.mark.parametrize("order", ["given", "hilbert"])
def test_ns_residual_norms(order):
    """gd_run's residuals are updateResidualSA's (evaluate2d.jl:6-48): per
    conservative variable, the sum over cells of (U_new - U_old)^2 of the
    run's last step, independent of the device order."""
    S = pkg().stage
    L = pkg()._lib
    mesh = _box("mixed", 30, 20)
    U = _box_state(mesh)
    s = S.GodunovSolver(mesh, _dev_params(), order=L.GD_ORDER_HILBERT if order == "hilbert" else L.GD_ORDER_AS_GIVEN)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    s.run(3, dt=1e-3)
    before = s.conservative()
    st = s.run(1, dt=1e-3, residuals=True)
    after = s.conservative()
    st2 = s.run(2, dt=1e-3, residuals=True)
    last = s.conservative()
    s.close()
    want = ((after - before) ** 2).sum(axis=0)
    got = np.array(st.residuals[:])
    assert np.all(want > 0)
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    # a longer run reports its last step only
    s = S.GodunovSolver(mesh, _dev_params())
    s.state.set_array(L.GD_ARR_U_OLD, U)
    s.run(5, dt=1e-3)
    prev = s.conservative()
    s.close()
    assert np.allclose(np.array(st2.residuals[:]), ((last - prev) ** 2).sum(axis=0), rtol=1e-12, atol=0)


def test_mesh_cell_order_reports_the_device_layout():
    """gd_mesh_cell_order: identity as given, the caller's map for a caller
    permutation, a permutation for RCM and Hilbert."""
    L = pkg()._lib
    M = pkg().mesh
    mesh = _box("mixed", 30, 20)
    n = mesh.nCells
    perm = M.random_permutation(n, SEED)
    for order, want in ((L.GD_ORDER_AS_GIVEN, np.arange(n)), (L.GD_ORDER_PERMUTATION, perm),
                        (L.GD_ORDER_RCM, None), (L.GD_ORDER_HILBERT, None)):
        dm = L.DeviceMesh(mesh, order, perm if order == L.GD_ORDER_PERMUTATION else None)
        got = dm.cell_order()
        dm.close()
        assert np.array_equal(np.sort(got), np.arange(n))
        if want is not None:
            assert np.array_equal(got, want)


@pytest.mark.parametrize("mixed", [False, True])
def test_ns_run_continues_where_the_last_ended(mixed):
    """A run that follows another on the same state starts from the primitives
    and dt inputs the last update left (no pass over U): run(3), run(0),
    run(4) ends bit for bit where run(7) does, wall force and flow time
    included, in FP64 and FP32-mixed.  Any other entry point in between makes
    the next run start from U afresh: after the state is read and written back
    the FP64 run is again bit-identical; after the caller writes U through
    gd_state_device_u, the run starts from what was written."""
    import ctypes as C
    S = pkg().stage
    L = pkg()._lib
    mesh = _box("mixed", 30, 20)
    U = _box_state(mesh)
    prm = _dev_params(mixed=mixed)

    def solver():
        s = S.GodunovSolver(mesh, prm, order=L.GD_ORDER_HILBERT)
        s.state.set_array(L.GD_ARR_U_OLD, U)
        return s

    kw = dict(cfl=0.4, dt_policy=L.GD_DT_LOCAL, wall_code=-3)
    a = solver()
    sa = a.run(7, **kw)
    ua = a.conservative()
    a.close()
    b = solver()
    b.run(3, **kw)
    b.run(0, **kw)
    sb = b.run(4, **kw)
    ub = b.conservative()
    b.close()
    assert np.array_equal(ua, ub)
    assert (sa.flow_time, sa.last_dt) == (sb.flow_time, sb.last_dt)
    assert tuple(sa.wall_force) == tuple(sb.wall_force) and sa.wall_force[0] != 0.0
    # the state read back and written again between the runs
    c = solver()
    c.run(3, **kw)
    c.state.set_array(L.GD_ARR_U_OLD, c.conservative())
    c.run(4, **kw)
    uc = c.conservative()
    c.close()
    if mixed:
        assert rel_linf(uc, ua) <= 1e-5
    else:
        assert np.array_equal(uc, ua)
    # the caller writes U on the device between two runs (Hilbert order on
    # the device: the caller's rows permuted by gd_mesh_cell_order)
    d = solver()
    d.run(3, **kw)
    U3 = d.conservative()
    W = np.asfortranarray(U3 * np.array([1.01, 1.0, 1.0, 1.02]))
    lib = L.load()
    Wd = np.asfortranarray(W[d.dmesh.cell_order()])
    ptr = C.POINTER(C.c_double)()
    assert lib.gd_state_device_u(d.state.handle, C.byref(ptr)) == 0
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    d.state.sync()
    assert rt.hipMemcpy(C.cast(ptr, C.c_void_p), Wd.ctypes.data, Wd.nbytes, 1) == 0  # host to device
    d.run(4, **kw)
    ud = d.conservative()
    d.close()
    e = S.GodunovSolver(mesh, prm, order=L.GD_ORDER_HILBERT)
    e.state.set_array(L.GD_ARR_U_OLD, W)
    e.flow_time = sb.flow_time  # not read by these boundaries; kept alike anyway
    e.run(4, **kw)
    ue = e.conservative()
    e.close()
    assert np.array_equal(ud, ue)
