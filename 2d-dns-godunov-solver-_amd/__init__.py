"""MI355X-native explicit Godunov step of dimaZloy/2D-DNS-godunov-solver-.

Hand-written HIP kernels for gfx950 behind the C ABI of include/godunov.h
(libgodunov.so), a ctypes binding (``_lib``), the reference's mesh and field
structs with mesh generators (``mesh``), mesh files in the reference's HDF5
format (``hdf5io``), node fields as VTK files (``vtkio``), and a mirror of
the reference's step interface (``stage``).  The directory name is not a Python identifier; import
it with ``importlib.import_module("2d-dns-godunov-solver-_amd")``.
"""
from . import _lib, hdf5io, mesh, stage, vtkio  # noqa: F401
from .hdf5io import read_mesh_hdf5, write_mesh_hdf5  # noqa: F401
from .vtkio import saveResults2VTK, saveResults4VTK, save_results_vtk  # noqa: F401
from .mesh import THERMOPHYSICS, fields2d, mesh2d_Int32  # noqa: F401
from .stage import (GodunovSolver, calcDiffTerm, calcOneStageCUDA, calcOneStageCUDAx32,  # noqa: F401
                    calcOneStageMUSCL, calcOneStageMUSCLcpuFast, cells2nodesSolutionReconstructionWithStencils,
                    configure, make_params, phs2dcns2dcellsSA, updateVariablesSA)

PACKAGE_NAME = "2d-dns-godunov-solver-_amd"
