"""Operator cores the engine compiles (generic/_laplacian.py, _biharmonic.py, _convection.py,
_source.py, dedicated/_navier_stokes.py:18-71, 221-280, dedicated/_ks_convection.py).

Each core carries an ``_engine_kind`` that register_mesh maps onto the engine's term
descriptor.  LinearCoef cores still return the reference's symbol tensors when called directly.
NonlinearFunc cores evaluated directly (outside an operator) run through libtfsm with the mask
switched off, as the reference's core call does not dealias by itself.
"""

from typing import Callable, Optional

import torch

from ..mesh import FourierMesh
from ._base import CoreGenerator, LinearCoef, LinearOperator, NonlinearFunc, NonlinearOperator, Operator


class _LaplacianCore(LinearCoef):
    _engine_kind = "lap"

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        return torch.cat([f_mesh.laplacian()] * n_channel, dim=1)


class Laplacian(LinearOperator):
    """nabla^2 of each channel (_laplacian.py:17-26)."""

    def __init__(self) -> None:
        super().__init__(_LaplacianCore())


class _BiharmonicCore(LinearCoef):
    _engine_kind = "biharm"

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        return torch.cat([f_mesh.laplacian() * f_mesh.laplacian()] * n_channel, dim=1)


class Biharmonic(LinearOperator):
    """nabla^4 of each channel (_biharmonic.py:20-29)."""

    def __init__(self) -> None:
        super().__init__(_BiharmonicCore())


class _SpatialDerivativeCore(LinearCoef):
    """(i a_dim)^order (_spatial_derivative.py:7-20); odd orders make L imaginary."""

    _engine_kind = "deriv"

    def __init__(self, dim_index: int, order: int) -> None:
        super().__init__()
        self.dim_index = dim_index
        self.order = order

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        return f_mesh.grad(self.dim_index, self.order)


class _SpatialDerivativeGenerator(CoreGenerator):
    def __init__(self, dim_index: int, order: int) -> None:
        self.dim_index = dim_index
        self.order = order

    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> LinearCoef:
        if n_channel != 1:
            raise ValueError("The SpatialDerivative operator only supports scalar field.")
        return _SpatialDerivativeCore(self.dim_index, self.order)


class SpatialDerivative(LinearOperator):
    """d^order / d x_dim^order of a scalar field (_spatial_derivative.py:40-55)."""

    def __init__(self, dim_index: int, order: int) -> None:
        super().__init__(_SpatialDerivativeGenerator(dim_index, order))


class _ImplicitUnitSourceCore(LinearCoef):
    _engine_kind = "unit"

    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        return torch.ones_like(f_mesh.bf_x)


class _ImplicitFuncSourceCore(NonlinearFunc):
    """f(u) source with an arbitrary python function: not an engine term."""

    def __init__(self, source_func: Callable[[torch.Tensor], torch.Tensor], non_linear: bool = True) -> None:
        super().__init__(non_linear)
        self.source_func = source_func

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        if u is None:
            u = f_mesh.ifft(u_fft).real
        return f_mesh.fft(self.source_func(u))


class ImplicitSource(Operator):
    """Implicit source (_source.py:44-74): unit form f(u) = u (a linear symbol), or a function."""

    def __init__(self, source_func: Optional[Callable[[torch.Tensor], torch.Tensor]] = None,
                 non_linear: bool = True) -> None:
        if source_func is None:
            generator = lambda f_mesh, n_channel: _ImplicitUnitSourceCore()  # noqa: E731
        else:
            generator = lambda f_mesh, n_channel: _ImplicitFuncSourceCore(source_func, non_linear)  # noqa: E731
        super().__init__(generator)


def _core_direct(kind, u_fft, f_mesh: FourierMesh):
    """Evaluate a product core on its raw input (no dealiasing) through the engine."""
    from ..engine import SpectralPlan, require_gpu_tensor

    require_gpu_tensor(u_fft, "u_fft")
    plan = SpectralPlan(f_mesh.mesh_info, u_fft.shape[0], u_fft.shape[1], nonlinear=kind, nonlinear_coef=1.0,
                        device=u_fft.device, dealias_rate=2.0)
    out = plan.apply(plan.full_to_half(u_fft))
    return plan.half_to_full(out)


class _ConvectionCore(NonlinearFunc):
    """(u . grad) u_c (_convection.py:9-48)."""

    _engine_kind = "conv"

    def __init__(self):
        super().__init__()

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _core_direct("conv", u_fft, f_mesh)


class _ConvectionGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int):
        if f_mesh.n_dim != n_channel:
            raise ValueError("convection operator only works for vector field with the same dimension as mesh")
        return _ConvectionCore()


class Convection(NonlinearOperator):
    """u . grad u for a vector field with C == D (_convection.py:67-76)."""

    def __init__(self) -> None:
        super().__init__(_ConvectionGenerator())


class _VorticityConvectionCore(NonlinearFunc):
    """(u . grad) omega with u = (d_y psi, -d_x psi), psi = -lap^-1 omega (_navier_stokes.py:18-46)."""

    _engine_kind = "vc"

    def __init__(self) -> None:
        super().__init__()

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _core_direct("vc", u_fft, f_mesh)


class _VorticityConvectionGenerator(CoreGenerator):
    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> NonlinearFunc:
        if f_mesh.n_dim != 2 or n_channel != 1:
            raise ValueError("Only vorticity in 2Dmesh is supported")
        return _VorticityConvectionCore()


class VorticityConvection(NonlinearOperator):
    """Vorticity advection in 2D (_navier_stokes.py:62-71)."""

    def __init__(self) -> None:
        super().__init__(_VorticityConvectionGenerator())


class _NSPressureConvectionCore(NonlinearFunc):
    """Velocity-form Navier-Stokes nonlinear term (_navier_stokes.py:221-262):
    p = lap^-1 (nabla . (conv - f)), N = nabla p - (conv - f) + f, conv = (u . grad) u.  With an
    external force the reference's core masks its input spectrum in place; the engine plan
    reproduces that (tfsm_plan_add_force).  `_engine_force` lists the force's explicit sources as
    (coefficient, spatial field)."""

    _engine_kind = "nsp"

    def __init__(self, force_sources=()) -> None:
        super().__init__()
        self._engine_force = list(force_sources)

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        if self._engine_force:
            raise ValueError("term not supported by engine: the forced NSPressureConvection core alone "
                             "(call or integrate the operator)")
        return _core_direct("nsp", u_fft, f_mesh)


def _force_sources(force, f_mesh: FourierMesh, n_channel: int):
    """The explicit sources an external-force operator is made of, as (coefficient, spatial field):
    the reference evaluates force(u_fft=..., return_in_fourier=True) (_navier_stokes.py:244-247),
    which for explicit sources is the sum of coef * fftn(source)."""
    force.register_mesh(f_mesh, n_channel)
    if force._linear_cores:
        raise ValueError("term not supported by engine: an external force with linear terms "
                         "(the engine takes forces made of explicit sources)")
    out = []
    for coef, core in force._nonlinear_funcs:
        if getattr(core, "_engine_kind", None) != "src":
            raise ValueError(f"term not supported by engine: an external force with {type(core).__name__} "
                             "(the engine takes forces made of explicit sources)")
        if isinstance(coef, torch.Tensor):
            if coef.numel() != 1:
                raise ValueError("term not supported by engine: per-sample coefficient on an external force")
            coef = float(coef)
        out.append((float(coef), core._source_spatial))
    return out


class _NSPressureConvectionGenerator(CoreGenerator):
    def __init__(self, external_force=None) -> None:
        self.external_force = external_force

    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> NonlinearFunc:
        if f_mesh.n_dim != n_channel or f_mesh.n_dim == 1:
            raise ValueError("term not supported by engine: NSPressureConvection needs a 2D or 3D velocity field "
                             "with one channel per dimension")
        if self.external_force is not None:
            return _NSPressureConvectionCore(_force_sources(self.external_force, f_mesh, n_channel))
        return _NSPressureConvectionCore()


class NSPressureConvection(NonlinearOperator):
    """-grad(lap^-1 div((u . grad) u - f)) - (u . grad) u + f (_navier_stokes.py:265-280); f is an
    operator made of explicit sources, or None."""

    def __init__(self, external_force=None) -> None:
        super().__init__(_NSPressureConvectionGenerator(external_force))


class _KSConvectionCore(NonlinearFunc):
    """1/2 |grad phi|^2 of a scalar field, minus its mean when remove_mean (_ks_convection.py:8-38).
    The reference's mean runs over the whole tensor, batch included, so with remove_mean the
    samples of a batch share it."""

    def __init__(self, remove_mean: bool) -> None:
        super().__init__()
        self.remove_mean = remove_mean
        self._engine_kind = "ks" if remove_mean else "ks_keep_mean"

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        return _core_direct(self._engine_kind, u_fft, f_mesh)


class _KSConvectionGenerator(CoreGenerator):
    def __init__(self, remove_mean: bool) -> None:
        self.remove_mean = remove_mean

    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> NonlinearFunc:
        if n_channel != 1:
            raise NotImplementedError("KSConvection only supports scalar field")
        return _KSConvectionCore(self.remove_mean)


class KSConvection(NonlinearOperator):
    """Kuramoto-Sivashinsky convection 1/2 |grad phi|^2 (_ks_convection.py:55-67)."""

    def __init__(self, remove_mean: bool = True) -> None:
        super().__init__(_KSConvectionGenerator(remove_mean))
