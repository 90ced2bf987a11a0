"""gd_calc_diff_term_noalloc (calcDiffTerm on caller-owned device arrays, its
intermediates in the library's static scratch array) against the state-based
gd_calc_diff_term, through the C ABI: bit for bit, on two calls in a row, the
second on a smaller mesh; the scratch bound and argument checks.""" 
import ctypes as C

import numpy as np
import pytest

from conftest import pkg
from helpers import SEED, tgv_fields

pytestmark = pytest.mark.gpu


def _case(mesh):
    S, L = pkg().stage, pkg()._lib
    prm = S.make_params(viscous=True, energy_law=L.GD_ENERGY_CONST, mu_const=1.0 / 1600, k_const=1.0 / 0.72,
                        RGAS=1.0, Gamma=1.4, bcs={})
    f = tgv_fields(mesh, noise=1e-3, seed=SEED)
    s = S.GodunovSolver(mesh, prm, order=L.GD_ORDER_HILBERT)
    s.set_primitives(f.densityCells, f.UxCells, f.UyCells, f.pressureCells)
    return s, prm, f


class _Hip:
    """Device buffers through the HIP runtime that libgodunov links
    (test plumbing: hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape):
        a = np.empty(shape)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self, *ps):
        for p in ps:
            assert self.rt.hipFree(p) == 0


def test_diff_term_noalloc_matches_the_state_entry_point():
    L = pkg()._lib
    lib = L.load()
    M = pkg().mesh
    for n in (96, 40):  # 18,432 then 3,200 triangles: the second call on a smaller mesh
        mesh = M.make_tgv_periodic_mesh(n)
        s, prm, f = _case(mesh)
        print(f"noalloc calcDiffTerm: {mesh.nCells} cells", flush=True)
        s.state.calc_diff_term()
        ref = s.state.get_array(L.GD_ARR_U_DIFF)       # caller order
        U = s.conservative()
        order = s.dmesh.cell_order()                    # device cell d holds caller cell order[d]
        q = np.stack([f.densityCells, f.UxCells, f.UyCells, f.pressureCells], 1)
        hip = _Hip()
        dU = hip.upload(U[order].T)                     # 4 x n_cells, device order
        dQ = hip.upload(q[order].T)
        dD = hip.upload(np.full((4, mesh.nCells), 3.0))
        rc = lib.gd_calc_diff_term_noalloc(s.dmesh.handle, C.byref(prm), dU, dQ, dD, None)
        L.check(rc)
        out = np.empty((mesh.nCells, 4))
        out[order] = hip.download(dD, (4, mesh.nCells)).T
        hip.free(dU, dQ, dD)
        s.close()
        assert np.all(out[:, 0] == 3.0)                 # column 0 left as it is
        assert np.array_equal(out[:, 1:], ref[:, 1:])
        assert np.abs(ref[:, 1:]).max() > 0


def test_diff_term_noalloc_scratch_bound_and_arguments():
    L = pkg()._lib
    lib = L.load()
    assert lib.gd_diff_term_scratch_bytes(2_000_000, 1_000_000) <= 256 << 20
    assert lib.gd_diff_term_scratch_bytes(4_000_000, 2_000_000) > 256 << 20
    M = pkg().mesh
    mesh = M.make_tgv_periodic_mesh(8)
    s, prm, _ = _case(mesh)
    rc = lib.gd_calc_diff_term_noalloc(s.dmesh.handle, C.byref(prm), None, None, None, None)
    assert rc == L.GD_ERR_INVALID
    s.close()
