"""Operators with the reference's names (torchfsm.operator)."""

from ._base import (
    CoreGenerator,
    ExplicitSource,
    LinearCoef,
    LinearOperator,
    NonlinearFunc,
    NonlinearOperator,
    Operator,
    OperatorLike,
    check_value_with_mesh,
)
from ._diag import (
    ConservativeConvection,
    Curl,
    Div,
    Grad,
    Velocity2Pressure,
    Vorticity2Pressure,
    Vorticity2Velocity,
)
from ._cores import (
    Biharmonic,
    Convection,
    ImplicitSource,
    KSConvection,
    Laplacian,
    NSPressureConvection,
    SpatialDerivative,
    VorticityConvection,
)

__all__ = [
    "Operator",
    "OperatorLike",
    "LinearOperator",
    "NonlinearOperator",
    "LinearCoef",
    "NonlinearFunc",
    "CoreGenerator",
    "ExplicitSource",
    "ImplicitSource",
    "Laplacian",
    "Biharmonic",
    "Convection",
    "VorticityConvection",
    "NSPressureConvection",
    "KSConvection",
    "SpatialDerivative",
    "Div",
    "Grad",
    "Curl",
    "Vorticity2Velocity",
    "Vorticity2Pressure",
    "Velocity2Pressure",
    "ConservativeConvection",
    "check_value_with_mesh",
]
