"""Function forms of the operators (functional.py:9-360): one engine evaluation each."""

from .operator import (
    Biharmonic,
    ConservativeConvection,
    Convection,
    Curl,
    Div,
    Grad,
    KSConvection,
    Laplacian,
    SpatialDerivative,
    Velocity2Pressure,
    Vorticity2Pressure,
    Vorticity2Velocity,
    VorticityConvection,
)


def biharmonic(u=None, u_fft=None, mesh=None):
    """functional.py:9-28."""
    return Biharmonic()(u=u, u_fft=u_fft, mesh=mesh)


def conservative_convection(u=None, u_fft=None, mesh=None):
    """functional.py:31-50."""
    return ConservativeConvection()(u=u, u_fft=u_fft, mesh=mesh)


def convection(u=None, u_fft=None, mesh=None):
    """functional.py:53-73."""
    return Convection()(u=u, u_fft=u_fft, mesh=mesh)


def curl(u=None, u_fft=None, mesh=None):
    """functional.py:76-96."""
    return Curl()(u=u, u_fft=u_fft, mesh=mesh)


def div(u=None, u_fft=None, mesh=None):
    """functional.py:99-119."""
    return Div()(u=u, u_fft=u_fft, mesh=mesh)


def grad(u=None, u_fft=None, mesh=None):
    """functional.py:122-142."""
    return Grad()(u=u, u_fft=u_fft, mesh=mesh)


def laplacian(u=None, u_fft=None, mesh=None):
    """functional.py:145-164."""
    return Laplacian()(u=u, u_fft=u_fft, mesh=mesh)


def spatial_derivative(dim_index: int, order: int, u=None, u_fft=None, mesh=None):
    """functional.py:167-190."""
    return SpatialDerivative(dim_index, order)(u=u, u_fft=u_fft, mesh=mesh)


def ks_convection(u=None, u_fft=None, mesh=None, remove_mean: bool = True):
    """functional.py:193-214."""
    return KSConvection(remove_mean)(u=u, u_fft=u_fft, mesh=mesh)


def vorticity_convection(u=None, u_fft=None, mesh=None):
    """functional.py:216-236."""
    return VorticityConvection()(u=u, u_fft=u_fft, mesh=mesh)


def vorticity2velocity(vorticity=None, vorticity_fft=None, mesh=None):
    """functional.py:270-290."""
    return Vorticity2Velocity()(u=vorticity, u_fft=vorticity_fft, mesh=mesh)


def velocity2pressure(velocity=None, velocity_fft=None, mesh=None, external_force=None):
    """functional.py:293-324: -lap^-1 div(Convection(u)), Convection dealiased.  Laplacian.solve of
    -div is the same per-mode product as Velocity2Pressure's -div * invlap, so this runs that
    operator."""
    return Velocity2Pressure(external_force)(u=velocity, u_fft=velocity_fft, mesh=mesh)


def vorticity2pressure(vorticity=None, vorticity_fft=None, mesh=None, external_force=None):
    """functional.py:327-360: Convection (dealiased) of the vorticity's velocity, then as
    velocity2pressure.  The mask commutes with Vorticity2Velocity's per-mode multiplier, so this is
    Vorticity2Pressure's evaluation."""
    return Vorticity2Pressure(external_force)(u=vorticity, u_fft=vorticity_fft, mesh=mesh)
