"""Golden fixtures (tests/golden/*.npz, written by make_golden.py): the
oracle reproduces them bit for bit (CPU), and the device matches them at the
parity tolerances (GPU)."""
import os

import numpy as np
import pytest

from conftest import ROOT, pkg
from golden.make_golden import ns_case, ref_case  # noqa: E402
from helpers import rel_linf

GOLD = os.path.join(ROOT, "tests", "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name)))


def test_oracle_reproduces_reference_scheme_fixture():
    g = _load("ref_step_small.npz")
    now = ref_case()
    for k, v in g.items():
        assert np.array_equal(now[k], v), k


def test_oracle_reproduces_ns_fixture():
    g = _load("ns_step_small.npz")
    now = ns_case()
    for k, v in g.items():
        assert np.array_equal(now[k], v), k


@pytest.mark.gpu
def test_device_matches_fixtures():
    import oracle as O
    from golden.make_golden import NS_BCS, NS_GAS, REF_BCS
    S = pkg().stage
    L = pkg()._lib
    M = pkg().mesh
    g = _load("ref_step_small.npz")
    th = M.THERMOPHYSICS(RGAS=1.0, Gamma=1.4)
    for name, mesh, bcs in (("tgv8", M.make_tgv_periodic_mesh(8), {}),
                            ("box", M.make_box_mesh(6, 5, 2.0, 1.0, quads="mixed"), REF_BCS)):
        U = g[f"{name}_U"]
        cfg = S.SchemeConfig(bcs=dict(bcs), energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72)
        f = M.fields2d.zeros(mesh.nCells, mesh.nNodes)
        f.densityCells[:] = U[:, 0]; f.UxCells[:] = U[:, 1] / U[:, 0]; f.UyCells[:] = U[:, 2] / U[:, 0]
        f.pressureCells[:] = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0])
        mv = O.MeshView(mesh)
        O.cells2nodes(mv, f)
        ct = M.distibuteCellsInThreadsSA(2, mesh.nCells)
        Ud = np.zeros((mesh.nCells, 4), order="F")
        S.calcDiffTerm(ct, ct, mesh, f, None, th, U, None, Ud, config=cfg)
        FL = np.zeros_like(Ud); UN = np.zeros_like(Ud)
        S.calcOneStageMUSCLcpuFast(ct, 1.0, 1e-3, 0.0, mesh, f, th, U, FL, Ud, UN, config=cfg)
        assert rel_linf(Ud[:, 1:], g[f"{name}_Udiff"][:, 1:]) <= 1e-12
        assert rel_linf(UN, g[f"{name}_Unew"]) <= 1e-12
    n = _load("ns_step_small.npz")
    mesh = M.make_box_mesh(6, 5, 2.0, 1.0, quads="mixed")
    for tag, prec in (("fp64", L.GD_FP64), ("mixed", L.GD_FP32_MIXED)):
        p = S.make_params(scheme=L.GD_SCHEME_NS, precision=prec, venkat_K=NS_GAS["venkat_K"], RGAS=NS_GAS["R"],
                          mu_ref=NS_GAS["mu_ref"], T_ref=NS_GAS["T_ref"], S_suth=NS_GAS["S"], Pr=NS_GAS["Pr"],
                          bcs=NS_BCS)
        s = S.GodunovSolver(mesh, p)
        s.state.set_array(L.GD_ARR_U_OLD, n["ns_U"])
        s.run(1, dt=2e-3)
        tol = 1e-12 if tag == "fp64" else 2e-5
        assert rel_linf(s.conservative() - n["ns_U"], n[f"ns_{tag}_Unew"] - n["ns_U"]) <= tol
        s.close()
