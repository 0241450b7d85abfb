"""MI355X fractional 4-step incompressible Navier-Stokes finite-volume solver.

The time step runs as hand-written HIP kernels (fvm_amd/csrc, built into
fvm_amd/libfvm_amd.so) behind the C ABI of include/fvm_capi.h.  This package holds the thin
Python binding (capi, solver), the grid / initial-condition / restart formats (grid) and the
Fortran ISO_C_BINDING driver sources (fortran/).
"""
from .grid import OGrid, make_case, WORKLOADS  # noqa: F401
