"""mps3d_amd: MI355X-native implicit field-solve path of mps3d.

Geometric multigrid (red-black Gauss-Seidel, restriction, prolongation),
GMRES and the drift-diffusion / Poisson stencil operators as hand-written
HIP kernels for gfx950 behind the C ABI of include/mps3d.h.
"""
from .solver import FieldSolver, from_device, ghosted, to_device  # noqa: F401

__all__ = ["FieldSolver", "ghosted", "to_device", "from_device"]
