"""MI355X implicit coupled pseudo-time step (ICSFoam hot path).

Host-side mirror of the reference's interfaces over the C ABI of
libicsf.so:

  fvMesh                 polyMesh upload + geometry + addressing
  dbnsCase               fields, thermo, fvSchemes/fvSolution, and the
                         dbnsFoam pseudo-time step on the GPU
  convectiveFlux(...)    convectiveFluxScheme::calcFlux
  coupled solve pieces   createJacobian / matrixMul / precondition /
                         solveForIncr (blockFvMatrix, coupledMatrix)
"""
from ._lib import (IcsfError, KeywordError, NoDeviceError, ZeroDiagonalError, lib)  # noqa: F401
from .mesh import (colour_cells, colour_major_order, mesh_geometry, prepare_case, rcm_order,  # noqa: F401
                   renumber_mesh, sweep_levels)
from .solver import dbnsCase, fvMesh  # noqa: F401

__all__ = ["fvMesh", "dbnsCase", "rcm_order", "colour_cells", "colour_major_order", "renumber_mesh",
           "sweep_levels", "mesh_geometry", "prepare_case", "IcsfError", "KeywordError",
           "ZeroDiagonalError", "NoDeviceError", "lib"]
