"""Fields of the reference: Kolmogorov forcing (field.py:128-148) and the initial-condition
generators diffused_noise / truncated_fourier_series (field.py:8-125), with their transforms and
the diffusion step on the engine.  Random draws use torch's generator on the target device, in the
reference's order, so a seeded call gives the same draws as the reference on the same device."""

from typing import Optional, Sequence, Union

import torch

from .mesh import FourierMesh, MeshGrid, mesh_shape
from .operator import ExplicitSource, ImplicitSource, Laplacian, Operator


def kolm_force(x: torch.Tensor, drag_coef: float = -0.1, k: float = 4.0, length_scale: float = 1.0) -> Operator:
    """Kolmogorov forcing in vorticity form: a w - k cos(k l x)  (field.py:128-148)."""
    return drag_coef * ImplicitSource() - ExplicitSource(k * torch.cos(k * length_scale * x))


def _normalise(u: torch.Tensor, zero_centered: bool, unit_variance: bool, unit_magnitude: bool) -> torch.Tensor:
    dims = [i for i in range(1, u.ndim)]
    if zero_centered:
        u = u - u.mean(dim=dims, keepdim=True)
    if unit_variance:
        u = u / u.std(dim=dims, keepdim=True)
    if unit_magnitude:
        peak = torch.max(u.abs().view(u.size(0), -1), dim=1).values
        u = u / peak.view([u.shape[0]] + [1] * (u.ndim - 1))
    return u


def _device_dtype(mesh, device, dtype):
    if device is None and isinstance(mesh, (FourierMesh, MeshGrid)):
        device = mesh.device
    if dtype is None and isinstance(mesh, (FourierMesh, MeshGrid)):
        dtype = mesh.dtype
    return device, dtype


def diffused_noise(mesh: Union[Sequence[tuple], MeshGrid, FourierMesh], diffusion_coef: float = 1.0,
                   zero_centered: bool = True, unit_variance: bool = False, unit_magnitude: bool = True,
                   device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                   batch_size: int = 1, n_channel: int = 1) -> torch.Tensor:
    """Gaussian noise diffused by one exact step (dt = 1) of diffusion_coef * Laplacian(), then
    normalised (field.py:8-63)."""
    if unit_magnitude and unit_variance:
        raise ValueError("unit_magnitude and unit_variance are mutually exclusive.")
    device, dtype = _device_dtype(mesh, device, dtype)
    u_0 = torch.randn(*mesh_shape(mesh, batch_size=batch_size, n_channel=n_channel), device=device, dtype=dtype)
    u_0 = (diffusion_coef * Laplacian()).integrate(u_0, dt=1, step=1, mesh=mesh)
    return _normalise(u_0, zero_centered, unit_variance, unit_magnitude)


def truncated_fourier_series(mesh: Union[Sequence[tuple], MeshGrid, FourierMesh], freq_threshold: int = 5,
                             amplitude_range: tuple = (-1.0, 1.0), angle_range: tuple = (0.0, 2.0 * torch.pi),
                             zero_centered: bool = True, unit_variance: bool = False, unit_magnitude: bool = True,
                             device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                             batch_size: int = 1, n_channel: int = 1) -> torch.Tensor:
    """Random Fourier coefficients a e^{i phi} kept for |f_i| <= freq_threshold on every axis,
    inverse transformed (real part) and normalised (field.py:65-125)."""
    if unit_magnitude and unit_variance:
        raise ValueError("Cannot set both unit_magnitude and unit_variance to True.")
    device, dtype = _device_dtype(mesh, device, dtype)
    if not isinstance(mesh, FourierMesh):
        mesh = FourierMesh(mesh, device=device, dtype=dtype)
    shape = mesh_shape(mesh, batch_size=batch_size, n_channel=n_channel)
    magnitude = torch.rand(*shape, device=device, dtype=dtype) * (amplitude_range[1] - amplitude_range[0]) \
        + amplitude_range[0]
    angle = torch.rand(*shape, device=device, dtype=dtype) * (angle_range[1] - angle_range[0]) + angle_range[0]
    noise = magnitude * torch.exp(1j * angle)
    noise = noise * mesh.abs_low_pass_filter(freq_threshold)
    mesh.abs_low_pass_filter.cache_clear()
    noise = mesh.ifft(noise).real
    return _normalise(noise, zero_centered, unit_variance, unit_magnitude)
