"""Equation factories of the reference (pde.py:13-83) that the engine integrates."""

from typing import Optional

from .operator import (
    Biharmonic,
    Convection,
    KSConvection,
    Laplacian,
    NSPressureConvection,
    Operator,
    SpatialDerivative,
    VorticityConvection,
)


def Burgers(nu) -> Operator:
    """u_t = -u . grad u + nu lap u  (pde.py:13-25)."""
    return nu * Laplacian() - Convection()


def KuramotoSivashinsky() -> Operator:
    """phi_t = -phi_xx - phi_xxxx - phi phi_x, 1D only (pde.py:27-39)."""
    ks_eqn = -Laplacian() - Biharmonic() - Convection()
    ks_eqn.register_additional_check(lambda dim_value, dim_mesh: dim_value == 1 and dim_mesh == 1)
    return ks_eqn


def KuramotoSivashinskyHighDim() -> Operator:
    """phi_t = -lap phi - lap^2 phi - 1/2 |grad phi|^2 in 1D, 2D or 3D (pde.py:41-49)."""
    return -Laplacian() - Biharmonic() - KSConvection()


def KortewegDeVries(dispersion_coef=1, convection_coef: float = 6.0) -> Operator:
    """phi_t = -c1 phi_xxx + c2 phi phi_x (pde.py:51-64): an imaginary linear symbol."""
    return -dispersion_coef * SpatialDerivative(0, 3) + convection_coef * Convection()


def NavierStokesVorticity(Re: float, force: Optional[Operator] = None) -> Operator:
    """w_t + (u . grad) w = lap w / Re + curl f  (pde.py:66-83)."""
    ns_vorticity = -VorticityConvection() + 1 / Re * Laplacian()
    if force is not None:
        ns_vorticity += force
    return ns_vorticity


def NavierStokes(Re: float, force: Optional[Operator] = None) -> Operator:
    """u_t = -grad p - (u . grad) u + lap u / Re (+ f), velocity form in 2D/3D (pde.py:85-100).
    The engine covers force = None; add a forcing term to the returned operator instead."""
    return NSPressureConvection(force) + 1 / Re * Laplacian()
