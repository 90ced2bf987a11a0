"""The tangent-linear ns step of the oracle (the ns step instantiated with
forward-mode dual numbers), which the discrete adjoint is checked against.

It is pinned here by what the mathematics fixes: its value part is the FP64
step itself, bit for bit; its derivative part is linear in the seed; and it
matches central differences of the FP64 step, for the state, the fixed
boundary states and the wall force, on a mixed triangle/quad box with every
boundary kind."""
import numpy as np
import pytest

import oracle as O
from conftest import pkg
from helpers import SEED

BCS = {-1: (O.BC_FIXED, (1.1, 0.8, 0.1, 1.2)), -2: (O.BC_EXTRAPOLATE, ()), -3: (O.BC_ISOTHERMAL_WALL, (1.05,)),
       -4: (O.BC_SLIP_WALL, ())}
GAS = dict(R=1.0, mu_ref=2e-3, T_ref=1.0, S=0.4, Pr=0.72, venkat_K=3.0, venkat_scale=(1.0, 1.0, 1.0, 1.0))
DT = 2e-3


def _setup(flux=O.FLUX_AUSM_PLUS, viscous=True):
    M = pkg().mesh
    mesh = M.make_box_mesh(10, 7, 2.0, 1.0, quads="mixed")
    n = mesh.nCells
    u = M.uniform01(SEED, 4 * n).reshape(4, n)
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    # smooth field plus a small seeded perturbation (keeps limiter switches rare)
    r = 1.0 + 0.2 * np.sin(2 * x) * np.cos(3 * y) + 0.01 * u[0]
    vx = 0.6 + 0.3 * np.cos(x + y) + 0.01 * u[1]
    vy = 0.1 * np.sin(3 * x) + 0.01 * u[2]
    p = 1.2 + 0.2 * np.cos(2 * x - y) + 0.01 * u[3]
    U = np.zeros((n, 4), order="F")
    U[:, 0] = r; U[:, 1] = r * vx; U[:, 2] = r * vy; U[:, 3] = p / 0.4 + 0.5 * r * (vx * vx + vy * vy)
    prm = O.ns_params(gamma=1.4, flux=flux, viscous=viscous, bcs=BCS, **GAS)
    return mesh, O.MeshView(mesh), prm, U


def _direction(n, seed):
    M = pkg().mesh
    d = M.normal(seed, 4 * n).reshape(n, 4, order="F")
    return np.asfortranarray(d)


@pytest.mark.parametrize("viscous", [False, True])
def test_tl_value_is_the_fp64_step_bitwise(viscous):
    mesh, mv, prm, U = _setup(viscous=viscous)
    UN, _, _, f = O.ns_step(mv, prm, DT, U)
    UNt, _, ft, _ = O.ns_step_tl(mv, prm, DT, U, np.zeros_like(U))
    assert np.array_equal(UN, UNt)
    assert np.array_equal(f, ft)


def test_tl_is_linear_in_the_seed():
    mesh, mv, prm, U = _setup()
    d1, d2 = _direction(mesh.nCells, 1), _direction(mesh.nCells, 2)
    _, a, _, fa = O.ns_step_tl(mv, prm, DT, U, d1)
    _, b, _, fb = O.ns_step_tl(mv, prm, DT, U, d2)
    _, c, _, fc = O.ns_step_tl(mv, prm, DT, U, np.asfortranarray(2.0 * d1 - 3.0 * d2))
    assert np.allclose(c, 2.0 * a - 3.0 * b, rtol=1e-12, atol=1e-12 * np.abs(c).max())
    assert np.allclose(fc, 2.0 * fa - 3.0 * fb, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("flux", [O.FLUX_AUSM_PLUS, O.FLUX_ROE])
@pytest.mark.parametrize("viscous", [False, True])
def test_tl_matches_central_differences_of_the_state(flux, viscous):
    mesh, mv, prm, U = _setup(flux, viscous)
    d = _direction(mesh.nCells, 3) * np.abs(U).max(axis=0)
    d = np.asfortranarray(d)
    _, dUN, _, df = O.ns_step_tl(mv, prm, DT, U, d)
    h = 1e-6
    Up, _, _, fp = O.ns_step(mv, prm, DT, np.asfortranarray(U + h * d))
    Um, _, _, fm = O.ns_step(mv, prm, DT, np.asfortranarray(U - h * d))
    fd = (Up - Um) / (2 * h)
    err = np.abs(fd - dUN).max(axis=0) / np.abs(dUN).max(axis=0)
    print("tl vs fd", flux, viscous, err)
    assert err.max() < 1e-6
    dff = (fp - fm) / (2 * h)
    assert np.allclose(df, dff, rtol=1e-5, atol=1e-6 * np.abs(df).max())


def test_tl_matches_central_differences_of_the_fixed_state():
    mesh, mv, prm, U = _setup()
    dbc = np.zeros((O.MAX_BC + 1, 4))
    dbc[1] = (0.3, -0.5, 0.7, 0.2)  # code -1, the fixed state
    _, dUN, _, df = O.ns_step_tl(mv, prm, DT, U, np.zeros_like(U), dbc=dbc)
    assert np.abs(dUN).max() > 0
    h = 1e-6
    outs = []
    for sgn in (1, -1):
        q = np.array(BCS[-1][1]) + sgn * h * dbc[1]
        bcs = {**BCS, -1: (O.BC_FIXED, tuple(q))}
        p2 = O.ns_params(gamma=1.4, bcs=bcs, **GAS)
        outs.append(O.ns_step(mv, p2, DT, U))
    fd = (outs[0][0] - outs[1][0]) / (2 * h)
    assert np.abs(fd - dUN).max() <= 1e-6 * np.abs(dUN).max()


def test_tl_rejects_mixed_precision():
    mesh, mv, _, U = _setup()
    prm = O.ns_params(gamma=1.4, mixed=True, bcs=BCS, **GAS)
    with pytest.raises(RuntimeError):
        O.ns_step_tl(mv, prm, DT, U, np.zeros_like(U))
