"""Operators that change the channel count or are evaluated once, never stepped: Div, Grad, Curl
(generic/_div.py, _grad.py, _curl.py), Vorticity2Velocity, Vorticity2Pressure, Velocity2Pressure
(dedicated/_navier_stokes.py:75-217) and ConservativeConvection (generic/_conservative_convection.py).

Their cores carry an ``_engine_map`` tag; OperatorLike.__call__ routes a single such core to
map_call, which runs the engine's transforms, the tfsm_spectral_map kernels and, where a product is
needed, the Convection evaluation or tfsm_pointwise_products.  Integrating them raises ValueError
("term not supported by engine"), as their output is not a state of the same shape; the exception
is ConservativeConvection, which is also an engine product term (1D, 2D).
"""

from typing import Optional

import torch

from .. import _lib
from ..engine import SpectralPlan
from ..mesh import FourierMesh
from ._base import CoreGenerator, LinearCoef, LinearOperator, NonlinearFunc, NonlinearOperator

_CACHE = {}


def _plan(f_mesh: FourierMesh, batch: int, channels: int, device, nonlinear=None, dealias_rate=2 / 3):
    key = (tuple(f_mesh.mesh_info), batch, channels, str(device), nonlinear, dealias_rate)
    plan = _CACHE.get(key)
    if plan is None:
        if len(_CACHE) > 32:
            _CACHE.clear()
        plan = SpectralPlan(f_mesh.mesh_info, batch, channels, nonlinear=nonlinear,
                            nonlinear_coef=1.0 if nonlinear else 0.0, device=device, dealias_rate=dealias_rate)
        _CACHE[key] = plan
    return plan


def _to_half(plan: SpectralPlan, u, u_fft):
    if u_fft is None:
        if u.dtype != torch.float32:
            raise TypeError(f"the MI355X engine computes in float32; got u of dtype {u.dtype}")
        return plan.forward(u)
    if u_fft.dtype != torch.complex64:
        raise TypeError(f"the MI355X engine computes in complex64; got u_fft of dtype {u_fft.dtype}")
    return plan.full_to_half(u_fft)


def map_call(core, coef, u, u_fft, f_mesh: FourierMesh, return_in_fourier: bool, dealias: bool = True):
    """coef * core(u) for a diagnostic core on the engine.  dealias: the closure's mask on the input
    of dealiased cores (off for a core called directly, as the reference's core does not mask)."""
    x = u if u is not None else u_fft
    B, C, dev, D = x.shape[0], x.shape[1], x.device, f_mesh.n_dim
    kind = core._engine_map
    if isinstance(coef, torch.Tensor) or not isinstance(coef, (int, float)):
        raise ValueError("term not supported by engine: a tensor coefficient on a channel-changing operator")
    c = float(coef)
    p_in = _plan(f_mesh, B, C, dev)
    xh = _to_half(p_in, u, u_fft)
    if kind == "div":
        out, cout = p_in.spectral_map(_lib.MAP_DIV, xh, 1, c), 1
    elif kind == "grad":
        out, cout = p_in.spectral_map(_lib.MAP_GRAD, xh, D, c), D
    elif kind == "curl":
        cout = 1 if D == 2 else 3
        out = p_in.spectral_map(_lib.MAP_CURL, xh, cout, c)
    elif kind == "v2v":
        out, cout = p_in.spectral_map(_lib.MAP_V2V, xh, 2, c), 2
    elif kind == "vort2p":
        # dealiased vorticity -> velocity -> Convection of it (no further mask) -> pressure
        wd = p_in.spectral_map(_lib.MAP_DEALIAS, xh, 1) if dealias else xh
        vel = p_in.spectral_map(_lib.MAP_V2V, wd, 2)
        conv = _plan(f_mesh, B, 2, dev, nonlinear="conv", dealias_rate=2.0).apply(vel)
        out, cout = _plan(f_mesh, B, 2, dev).spectral_map(_lib.MAP_PRESSURE, conv, 1, c), 1
    elif kind == "vel2p":
        conv = _plan(f_mesh, B, C, dev, nonlinear="conv", dealias_rate=2 / 3 if dealias else 2.0).apply(xh)
        out, cout = p_in.spectral_map(_lib.MAP_PRESSURE, conv, 1, c), 1
    elif kind == "cons":
        ud = p_in.spectral_map(_lib.MAP_DEALIAS, xh, C) if dealias else xh
        phys = p_in.inverse(ud)
        uu = p_in.pointwise_products(phys)
        p_uu = _plan(f_mesh, B, C * C, dev)
        out, cout = p_in.spectral_map(_lib.MAP_CONSDIV, p_uu.forward(uu), C, c), C
    else:
        raise ValueError(f"term not supported by engine: {kind}")
    p_out = _plan(f_mesh, B, cout, dev)
    if return_in_fourier:
        return p_out.half_to_full(out)
    return p_out.inverse(out)


def _direct(core, u_fft, f_mesh):
    return map_call(core, 1.0, None, u_fft, f_mesh, True, dealias=False)


class _DivCore(NonlinearFunc):
    """sum_i d_i u_i (_div.py:9-28), no dealiasing."""

    _engine_map = "div"

    def __init__(self):
        super().__init__(False)

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _direct(self, u_fft, f_mesh)


class _DivGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != n_channel:
            raise ValueError("div operator only works for vector field with the same dimension as mesh")
        return _DivCore()


class Div(NonlinearOperator):
    """Divergence of a vector field with C == D (_div.py:48-60)."""

    def __init__(self) -> None:
        super().__init__(_DivGenerator())


class _GradCore(LinearCoef):
    """(i a_i) per axis (_grad.py:9-15); called directly it returns the symbol, as the reference's."""

    _engine_map = "grad"

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        return f_mesh.nabla_vector(1)


class _GradGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> LinearCoef:
        if n_channel != 1:
            raise ValueError("The Grad operator only supports scalar field.")
        return _GradCore()


class Grad(LinearOperator):
    """Gradient of a scalar field (_grad.py:30-42)."""

    def __init__(self) -> None:
        super().__init__(_GradGenerator())


class _CurlCore(NonlinearFunc):
    """Curl of a 2D (-> scalar) or 3D (-> vector) field (_curl.py:9-56), no dealiasing."""

    _engine_map = "curl"

    def __init__(self):
        super().__init__(False)

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _direct(self, u_fft, f_mesh)


class _CurlGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != n_channel:
            raise ValueError("div operator only works for vector field with the same dimension as mesh")
        if n_channel > 3 or n_channel < 2:
            raise ValueError("div operator only works for 2D or 3D vector field")
        return _CurlCore()


class Curl(NonlinearOperator):
    """Curl (_curl.py:79-90)."""

    def __init__(self) -> None:
        super().__init__(_CurlGenerator())


class _Vorticity2VelocityCore(LinearCoef):
    """u = (d_y psi, -d_x psi), psi = -lap^-1 w (_navier_stokes.py:75-91); called directly it
    returns the symbol pair, as the reference's."""

    _engine_map = "v2v"

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        n0, n1 = f_mesh.mesh_info[0][-1], f_mesh.mesh_info[1][-1]
        return -1 * f_mesh.invert_laplacian() * torch.cat(
            [f_mesh.grad(1, 1).repeat([1, 1, n0, 1]), -f_mesh.grad(0, 1).repeat([1, 1, 1, n1])], dim=1)


class _Vorticity2VelocityGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != 2 or n_channel != 1:
            raise ValueError("Only vorticity in 2Dmesh is supported")
        return _Vorticity2VelocityCore()


class Vorticity2Velocity(LinearOperator):
    """Velocity of a 2D vorticity field (_navier_stokes.py:106-115)."""

    def __init__(self) -> None:
        super().__init__(_Vorticity2VelocityGenerator())


def _no_force(external_force, name):
    if external_force is not None:
        raise ValueError(f"term not supported by engine: {name} with an external force")


class _Vorticity2PressureCore(NonlinearFunc):
    """p = -lap^-1 div((u . grad) u) of the velocity of a 2D vorticity (_navier_stokes.py:119-144)."""

    _engine_map = "vort2p"

    def __init__(self, external_force=None) -> None:
        _no_force(external_force, "Vorticity2Pressure")
        super().__init__(True)

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _direct(self, u_fft, f_mesh)


class _Vorticity2PressureGenerator(CoreGenerator):
    def __init__(self, external_force=None) -> None:
        self.external_force = external_force

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != 2 or n_channel != 1:
            raise ValueError("Only vorticity in 2Dmesh is supported")
        return _Vorticity2PressureCore(self.external_force)


class Vorticity2Pressure(NonlinearOperator):
    """Pressure of a 2D vorticity field (_navier_stokes.py:162-170); external_force must be None."""

    def __init__(self, external_force=None) -> None:
        _no_force(external_force, "Vorticity2Pressure")
        super().__init__(_Vorticity2PressureGenerator(external_force))


class _Velocity2PressureCore(NonlinearFunc):
    """p = -lap^-1 div((u . grad) u) of a velocity field (_navier_stokes.py:174-206)."""

    _engine_map = "vel2p"

    def __init__(self, external_force=None) -> None:
        _no_force(external_force, "Velocity2Pressure")
        super().__init__(True)

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _direct(self, u_fft, f_mesh)


class Velocity2Pressure(NonlinearOperator):
    """Pressure of a velocity field (_navier_stokes.py:209-217); external_force must be None."""

    def __init__(self, external_force=None) -> None:
        super().__init__(_Velocity2PressureCore(external_force))


class _ConservativeConvectionCore(NonlinearFunc):
    """div(u u^T) of a vector field (_conservative_convection.py:9-27), dealiased input.  Stepped by
    the engine in 1D and 2D (a product term like Convection); a single application also runs
    through the spectral maps, in any dimension."""

    _engine_map = "cons"
    _engine_kind = "cons"

    def __init__(self) -> None:
        super().__init__()

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _direct(self, u_fft, f_mesh)


class _ConservativeConvectionGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != n_channel:
            raise ValueError("convection operator only works for vector field with the same dimension as mesh")
        return _ConservativeConvectionCore()


class ConservativeConvection(NonlinearOperator):
    """Conservative-form convection div(u u^T) (_conservative_convection.py:43-70)."""

    def __init__(self) -> None:
        super().__init__(_ConservativeConvectionGenerator())
