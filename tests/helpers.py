"""Shared set-ups for the tests: meshes, seeded states and comparisons."""
import math

import numpy as np

from conftest import pkg

SEED = 20261203


def M():
    return pkg().mesh


def rel_linf(a, b):
    """max |a - b| / max |b| per column (columns of (n, k) arrays)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]; b = b[:, None]
    den = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return (np.abs(a - b).max(axis=0) / den).max()


def tgv_fields(mesh, gamma=1.4, ma=0.1, noise=1e-6, seed=SEED):
    """Nondimensional TGV: u = sin x cos y, v = -cos x sin y, rho = 1 + N(0, noise),
    p = 1/(gamma Ma^2) + (cos 2x + cos 2y)/4 at cell centroids."""
    mm = M()
    n = mesh.nCells
    x, y = mesh.cell_mid_points[:, 0], mesh.cell_mid_points[:, 1]
    f = mm.fields2d.zeros(n, mesh.nNodes)
    f.densityCells[:] = 1.0 + noise * mm.normal(seed, n)
    f.UxCells[:] = np.sin(x) * np.cos(y)
    f.UyCells[:] = -np.cos(x) * np.sin(y)
    f.pressureCells[:] = 1.0 / (gamma * ma * ma) + (np.cos(2 * x) + np.cos(2 * y)) / 4.0
    return f


def random_fields(mesh, seed=SEED, rho=(0.5, 2.0), vel=(-1.0, 1.0), p=(0.5, 3.0)):
    mm = M()
    n = mesh.nCells
    u = mm.uniform01(seed, 4 * n).reshape(4, n)
    f = mm.fields2d.zeros(n, mesh.nNodes)
    f.densityCells[:] = rho[0] + (rho[1] - rho[0]) * u[0]
    f.UxCells[:] = vel[0] + (vel[1] - vel[0]) * u[1]
    f.UyCells[:] = vel[0] + (vel[1] - vel[0]) * u[2]
    f.pressureCells[:] = p[0] + (p[1] - p[0]) * u[3]
    return f


def conservative(f, gamma):
    r, u, v, p = f.densityCells, f.UxCells, f.UyCells, f.pressureCells
    U = np.zeros((len(r), 4), order="F")
    U[:, 0] = r; U[:, 1] = r * u; U[:, 2] = r * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return U
