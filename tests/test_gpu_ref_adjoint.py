"""Discrete adjoint of the reference scheme's loop (gd_adjoint_run with the
reference scheme) on the GPU.

J = <w, U_N> (+ the summed wall force of one boundary code) after N fixed-dt
steps of the reference loop (calcDiffTerm with constant mu and k, the fast
MUSCL stage with AUSM+, updateVariablesSA, the node interpolation) on the
periodic Taylor-Green mesh and on boxes of triangles and quads with every
boundary ghost.  For any direction d,
<dJ/dU0, d> must equal the derivative of J along d that the oracle's
tangent-linear reference loop carries in dual numbers (itself pinned to the
oracle loop bit for bit and to central differences): FP64, 1e-9 relative.
The state gets seeded noise on every primitive so that no limiter argument
sits exactly on a kink (the symmetric TGV has exact ties along its symmetry
lines, where one-sided derivatives differ).
"""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED, tgv_fields

pytestmark = pytest.mark.gpu

DT = 2e-3
MU, KC = 1.0 / 1600, 1.0 / 0.72


def _case(n=100, noise=1e-2):
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(n)
    th = M.THERMOPHYSICS(RGAS=1.0, Gamma=1.4)
    f = tgv_fields(mesh, noise=noise)
    r = M.uniform01(SEED + 3, 3 * mesh.nCells).reshape(3, -1) - 0.5
    f.UxCells[:] += noise * r[0]
    f.UyCells[:] += noise * r[1]
    f.pressureCells[:] *= 1.0 + noise * r[2]
    U = O.phs2dcns(f, th.Gamma, mesh.nCells)
    return mesh, O.MeshView(mesh), th, U


def _solver(mesh, limiter, viscous, U, flux=None, energy=None):
    S = pkg().stage
    L = pkg()._lib
    p = S.make_params(scheme=L.GD_SCHEME_REF, flux=L.GD_FLUX_AUSM_PLUS if flux is None else flux, limiter=limiter,
                      viscous=viscous, energy_law=L.GD_ENERGY_CONST if energy is None else energy, mu_const=MU,
                      k_const=KC, RGAS=1.0, Gamma=1.4, bcs={})
    s = S.GodunovSolver(mesh, p)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    return s


def _direction(mesh, U, seed):
    M = pkg().mesh
    d = M.normal(seed, 4 * mesh.nCells).reshape(mesh.nCells, 4, order="F") * np.abs(U).max(axis=0)
    return np.asfortranarray(d)


LIMS = {"harmonic": O.LIM_REF_HARMONIC, "vanAlbada": O.LIM_VAN_ALBADA_A, "minmod": O.LIM_MINMOD}


@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("lim", list(LIMS))
def test_ref_adjoint_dot_product_with_tangent_linear_oracle(lim, viscous):
    M = pkg().mesh
    mesh, mv, th, U = _case()
    n = mesh.nCells
    N = 6
    w = np.asfortranarray(M.normal(SEED + 7, 4 * n).reshape(n, 4, order="F"))
    s = _solver(mesh, LIMS[lim], viscous, U)
    g0, gbc, st = s.adjoint(N, DT, checkpoint_every=4, dJ_dUN=w)
    UN_dev = s.conservative()
    s.close()
    prm = O.params(limiter=LIMS[lim], include_udiff=True, energy_law=O.ENERGY_CONST, mu_const=MU, k_const=KC)
    errs = []
    for k in range(2):
        d = _direction(mesh, U, SEED + 20 + k)
        UN, dUN = O.ref_loop_tl(mv, th, prm, U, d, DT, N, viscous=viscous)
        tl = float((w * dUN).sum())
        adj = float((g0 * d).sum())
        errs.append(abs(adj - tl) / abs(tl))
        print(lim, viscous, "<dJ/dU0, d>", adj, "tangent", tl, "rel", errs[-1])
    J = float((w * UN).sum())
    print("J", st.J, J, "recomputed", st.recomputed_steps, "checkpoints", st.n_checkpoints,
          "ms fwd/rev", st.ms_forward, st.ms_reverse)
    assert max(errs) <= 1e-9
    assert abs(st.J - J) <= 1e-12 * abs(J)
    assert np.abs(UN_dev - UN).max() <= 1e-12 * np.abs(UN).max()
    assert st.recomputed_steps == 3 + 1 and st.n_checkpoints == 2
    assert not gbc.any()


def test_ref_adjoint_independent_of_checkpointing():
    """Every checkpoint interval recomputes the same states, and states kept
    from the forward sweep (record_budget) are the same states: the steps are
    transposed in the same order, with identical bits."""
    mesh, _, _, U = _case(60)
    seg_bytes = 32 * (mesh.nCells + mesh.nNodes)  # one step's Q and node means Qn
    out = []
    # 7 steps; a budget of k segments keeps the trailing k segments' states
    for c, budget, recomputed in ((1, 0, 0), (3, 0, 4), (7, 0, 6), (3, 6 * seg_bytes, 2), (3, 1 << 40, 0),
                                  (2, 4 * seg_bytes, 2)):
        s = _solver(mesh, O.LIM_REF_HARMONIC, True, U)
        g0, _, st = s.adjoint(7, DT, checkpoint_every=c, dJ_dUN=U, record_budget=budget)
        out.append((g0, st.J, s.conservative()))
        s.close()
        print(c, budget, "recomputed", st.recomputed_steps, "storage", st.storage_bytes)
        assert st.recomputed_steps == recomputed
    for g0, J, UN in out[1:]:
        assert np.array_equal(g0, out[0][0]) and J == out[0][1] and np.array_equal(UN, out[0][2])


def test_ref_adjoint_leaves_the_loop_state():
    """After the adjoint the state holds U_N with its primitives and node
    primitives, as after gd_run: continuing with gd_run gives the oracle
    loop's U_{N+K}."""
    mesh, mv, th, U = _case(60)
    s = _solver(mesh, O.LIM_VAN_ALBADA_A, True, U)
    s.adjoint(4, DT, checkpoint_every=2)
    s.run(3, dt=DT)
    UN = s.conservative()
    s.close()
    prm = O.params(limiter=O.LIM_VAN_ALBADA_A, include_udiff=True, energy_law=O.ENERGY_CONST, mu_const=MU, k_const=KC)
    ref, _ = O.ref_loop_tl(mv, th, prm, U, None, DT, 7)
    assert np.abs(UN - ref).max() <= 1e-12 * np.abs(ref).max()


def test_ref_adjoint_refuses_what_it_does_not_cover():
    L = pkg()._lib
    mesh, _, _, U = _case(20)
    for kw, msg in ((dict(limiter=O.LIM_VAN_LEER_A), "limiters"), (dict(limiter=0, flux=L.GD_FLUX_ROE), "AUSM"),
                    (dict(limiter=0, energy=L.GD_ENERGY_AIR_POLY), "constant mu")):
        lim = kw.pop("limiter")
        s = _solver(mesh, lim, True, U, **kw)
        with pytest.raises(L.GodunovError, match=msg):
            s.adjoint(2, DT)
        s.close()


def test_ref_adjoint_full_size_against_central_differences():
    """The 2,000,000-triangle TGV square (noisy state), 20 steps at dt 5e-5,
    J = <w, U_N>: <dJ/dU0, d> in three smooth seeded directions (a plane wave
    per variable, |k| in [1, 4]) against central differences of J from the
    GPU's own forward loop (the forward sweep of an adjoint call, which starts
    from the primitives of U0 as the adjoint does) at h = 1e-8, and the
    gradient unchanged between checkpoint intervals 5 and 20.

    J(U0 + h d) - J(U0 - h d) is summed as <w, U_N+ - U_N-> (a sum of 8M
    products accumulated one by one carries rounding of order 1e-8, as large
    as the difference itself at h = 1e-9).  The noisy state has limiter
    arguments near zero in many cells, where the limiter has a kink: at
    h = 1e-6 some are crossed and the differences are 1e-4 to 1e-3 off (at
    1e-8: 1.6e-6 to 6.3e-6)."""
    M = pkg().mesh
    L = pkg()._lib
    mesh, _, _, U = _case(1000)
    n = mesh.nCells
    N, dt = 20, 5e-5
    w = np.asfortranarray(M.normal(SEED + 9, 4 * n).reshape(n, 4, order="F"))
    s = _solver(mesh, O.LIM_REF_HARMONIC, True, U)
    g0, _, st = s.adjoint(N, dt, checkpoint_every=5, dJ_dUN=w)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    g1, _, _ = s.adjoint(N, dt, checkpoint_every=20, dJ_dUN=w)
    assert np.array_equal(g0, g1)

    def UN(U0):
        s.state.set_array(L.GD_ARR_U_OLD, U0)
        s.adjoint(N, dt, checkpoint_every=N, to_host=False)
        return s.conservative()

    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    scale = np.abs(U).max(axis=0)
    h = 1e-8
    errs = []
    for k in range(3):
        r = M.uniform01(SEED + 40 + k, 12).reshape(4, 3)
        d = np.asfortranarray(np.stack(
            [scale[v] * np.sin((1.0 + 3.0 * r[v, 0]) * (np.cos(6.3 * r[v, 1]) * x + np.sin(6.3 * r[v, 1]) * y)
                               + 6.3 * r[v, 2]) for v in range(4)], 1))
        adj = float((g0 * d).sum())
        fd = float((w * (UN(np.asfortranarray(U + h * d)) - UN(np.asfortranarray(U - h * d)))).sum()) / (2 * h)
        errs.append(abs(adj - fd) / abs(adj))
        print("direction", k, "adjoint", adj, "central difference", fd, "rel", errs[-1])
    s.close()
    assert max(errs) <= 1e-4


# boxes of triangles and quads with a fixed inflow (-1), zero gradient (-2), a
# slip wall (-3) and an isothermal wall (-4); a TGV box with mirror ghosts
BOX = {-1: ("FIXED", (1.3, 0.7, -0.2, 1.9)), -2: ("EXTRAPOLATE", ()), -3: ("SLIP_WALL", ()),
       -4: ("ISOTHERMAL_WALL", (1.1,))}
TGV_BOX = {c: ("TGV_MIRROR", (71.43, 1.0, 1.0, 0.05)) for c in (-1, -2, -3, -4)}


def _box(quads, tgv, size=None):
    from helpers import random_fields
    M = pkg().mesh
    if tgv:
        mesh = M.make_box_mesh(40, 34, 6.283185307179586, 6.283185307179586, quads=quads)
        f = tgv_fields(mesh, noise=1e-2)
        r = M.uniform01(SEED + 5, 3 * mesh.nCells).reshape(3, -1) - 0.5
        f.UxCells[:] += 1e-2 * r[0]
        f.UyCells[:] += 1e-2 * r[1]
        f.pressureCells[:] *= 1.0 + 1e-2 * r[2]
    else:
        nx, ny = size or (40, 26)
        mesh = M.make_box_mesh(nx, ny, 2.0, 1.0, quads=quads)
        f = random_fields(mesh, seed=SEED + 6)
    th = M.THERMOPHYSICS(RGAS=1.0, Gamma=1.4)
    return mesh, O.MeshView(mesh), th, O.phs2dcns(f, th.Gamma, mesh.nCells)


@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("quads,tgv,lim,size", [("mixed", False, "harmonic", None), ("all", False, "vanAlbada", None),
                                                ("none", False, "minmod", None), ("mixed", True, "harmonic", None),
                                                ("mixed", False, "minmod", (3, 2)), ("none", False, "harmonic", (2, 1))])
def test_ref_adjoint_boxes_against_tangent_linear_oracle(quads, tgv, lim, size, viscous):
    """<dJ/dU0, d> + <dJ/dbc, dbc> against the tangent-linear loop's
    derivative along (d, dbc), J = <w, U_N> + 0.7 Fx - 0.4 Fy over 5 steps
    from flow time 0.1 (F the force through the slip wall, or through code -2
    of the TGV box), checkpoints every 2.  The small boxes have fewer cells,
    nodes and face entries than one wavefront, so every kernel's last wave
    has lanes past the end."""
    S = pkg().stage
    L = pkg()._lib
    M = pkg().mesh
    mesh, mv, th, U = _box(quads, tgv, size)
    n = mesh.nCells
    bcs = TGV_BOX if tgv else BOX
    wc = -2 if tgv else -3
    N, dt, t0 = 5, 1e-3, 0.1
    fw = (0.7, -0.4)
    p = S.make_params(scheme=L.GD_SCHEME_REF, limiter=LIMS[lim], viscous=viscous, energy_law=L.GD_ENERGY_CONST,
                      mu_const=MU, k_const=KC, RGAS=1.0, Gamma=1.4,
                      bcs={c: (getattr(L, "GD_BC_" + k), d) for c, (k, d) in bcs.items()})
    w = np.asfortranarray(M.normal(SEED + 7, 4 * n).reshape(n, 4, order="F"))
    s = S.GodunovSolver(mesh, p, order=L.GD_ORDER_HILBERT)
    s.state.set_array(L.GD_ARR_U_OLD, U)
    s.flow_time = t0
    g0, gbc, st = s.adjoint(N, dt, checkpoint_every=2, wall_code=wc, force_weight=fw, dJ_dUN=w)
    s.close()
    oprm = O.params(limiter=LIMS[lim], include_udiff=True, energy_law=O.ENERGY_CONST, mu_const=MU, k_const=KC,
                    bcs={c: (getattr(O, "BC_" + k), d) for c, (k, d) in bcs.items()})
    errs = []
    for k in range(2):
        d = np.asfortranarray(M.normal(SEED + 30 + k, 4 * n).reshape(n, 4, order="F") * np.abs(U).max(axis=0))
        db = np.zeros((O.MAX_BC + 1, 4))
        if not tgv:
            db[1] = M.normal(SEED + 40 + k, 4)
        UN, dUN, F, dF = O.ref_loop_tl_force(mv, th, oprm, U, d, dt, N, wc, viscous, flow_time=t0, dbc=db)
        tl = float((w * dUN).sum()) + fw[0] * dF[0] + fw[1] * dF[1]
        adj = float((g0 * d).sum()) + float((gbc * db).sum())
        errs.append(abs(adj - tl) / abs(tl))
        print(quads, tgv, lim, viscous, "adjoint", adj, "tangent", tl, "rel", errs[-1])
    J = float((w * UN).sum()) + fw[0] * F[0] + fw[1] * F[1]
    print("J", st.J, J, "force", st.force_sum[0], st.force_sum[1], F)
    assert max(errs) <= 1e-9
    assert abs(st.J - J) <= 1e-11 * max(abs(J), float(np.abs(w * UN).sum()))
    assert np.allclose([st.force_sum[0], st.force_sum[1]], F, rtol=1e-11, atol=1e-13 * np.abs(F).max())
    if tgv:
        assert not gbc.any()
    else:
        assert np.abs(gbc[1]).max() > 0 and not gbc[2:].any()

