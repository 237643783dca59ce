"""prism_amd -- MI355X (gfx950) finite-volume assembly + BiCGSTAB path of Prism.

The compute path is libprism_hip.so (hand-written HIP kernels behind the C ABI in
include/prism_hip.h); this package is a thin Python mirror of the reference's API on top of
it.  There is no CPU fallback: using the package without the built library raises.
"""
from ._capi import BoundaryConditionError, InvalidMeshError, PrismError
from .core import (Backend, adjoint, Matrix, Mesh, MeshData, Momentum, Scalar, SolverResult, Tensor, Transport,
                   Vector, Velocity, algo, box, convection, courant_number, diffusion, hip_lib,
                   needs_pressure_reference, read_unv,
                   solver, source, temporal)
from . import output, turbulence

__all__ = [
    "Backend", "adjoint", "Matrix", "Mesh", "MeshData", "Momentum", "Scalar", "SolverResult", "Tensor", "Transport",
    "Vector", "Velocity", "algo", "box", "convection", "courant_number", "diffusion", "hip_lib",
    "needs_pressure_reference", "read_unv",
    "solver", "source", "temporal", "output", "turbulence", "PrismError", "BoundaryConditionError", "InvalidMeshError",
]
