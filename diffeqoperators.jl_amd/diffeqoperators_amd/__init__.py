"""diffeqoperators_amd -- DiffEqOperators.jl's operator-application path on MI355X.

The constructors and composition rules mirror the Julia package (CenteredDifference,
UpwindDifference, RobinBC / DirichletBC / NeumannBC, compose, L*Q, L1+L2, mul!); applying
an operator runs hand-written gfx950 HIP kernels in libdop.so (C ABI: include/dop.h), and so do
the concretizations `Array` / `sparse` and the direct solve `ldiv` (Julia's `A \\ u`).

    import torch
    from diffeqoperators_amd import CenteredDifference, Dirichlet0BC
    Δ = CenteredDifference(2, 2, h, 100)
    bc = Dirichlet0BC(float)
    du = Δ * bc * u          # u: torch.float64 tensor on the GPU

Arrays are C-ordered tensors whose shape is the Julia size reversed: Julia axis N is
tensor dimension -N.
"""
from . import _abi
from .apply import Plan, julia_size, mul_, plan_for_shape, plans_for
from .batch import Batch, RK4
from .serve import Server
from .bcs import (AffineBC, BoundaryPadded, ComposedMultiDimBC, Dirichlet0BC, DirichletBC, GeneralBC,
                  MultiDimBC, MultiDimDirectionalBC, Neumann0BC, NeumannBC, PeriodicBC,
                  RobinBC, compose, decompose)
from .composite import DiffEqOperatorCombination, GhostDerivativeOperator
from .concretize import Array, Solver, ldiv, sparse
from .integrate import Tsit5
from .vector_calculus import Curl, Divergence, Gradient
from .nonlinear import nonlinear_diffusion, nonlinear_diffusion_
from .fornberg import calculate_weights
from .operators import CenteredDifference, DerivativeOperator, UpwindDifference

mul = mul_

__all__ = [
    "CenteredDifference", "UpwindDifference", "DerivativeOperator", "calculate_weights",
    "RobinBC", "DirichletBC", "Dirichlet0BC", "NeumannBC", "Neumann0BC", "MultiDimBC",
    "MultiDimDirectionalBC", "ComposedMultiDimBC", "AffineBC", "PeriodicBC", "GeneralBC",
    "compose", "decompose", "BoundaryPadded", "GhostDerivativeOperator", "DiffEqOperatorCombination",
    "mul_", "mul", "plans_for", "plan_for_shape", "Plan", "Batch", "RK4", "Server",
    "julia_size", "Array", "sparse", "ldiv", "Solver", "Tsit5", "Gradient", "Divergence", "Curl",
    "nonlinear_diffusion", "nonlinear_diffusion_",
]


def library_path() -> str:
    return _abi.LIB_PATH
