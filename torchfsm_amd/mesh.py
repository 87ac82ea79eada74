"""MeshGrid / FourierMesh with the reference's interface (mesh.py:9-527).

Coordinates, frequencies and symbols are built exactly as the reference builds them (they are
setup data: small broadcast tensors).  The transforms FourierMesh.fft / .ifft run on the MI355X
through libtfsm (hand-written Stockham kernels); there is no CPU transform path.
"""

from functools import lru_cache
from typing import Optional, Sequence, Tuple, Union

import numpy as np
import torch

from .utils import default, format_device_dtype


class MeshGrid:
    """Mesh coordinates of a periodic box (mesh.py:9-146).  x_i = (end - start) * i / n (start is
    not added, as in the reference)."""

    def __init__(self, mesh_info: Sequence[Tuple[float, float, int]], device=None, dtype=None) -> None:
        for dim_i in mesh_info:
            if len(dim_i) != 3:
                raise ValueError("each dimension should be a tuple of (start,end,n_points)")
        self.mesh_info = mesh_info
        self.meshs = [[] for _ in range(len(mesh_info))]
        self._dim_names = ["x", "y", "z"]
        self.device, self.dtype = format_device_dtype(device, dtype)
        self.n_dim = len(mesh_info)

    def __len__(self):
        return len(self.mesh_info)

    def __getitem__(self, idx: int):
        if len(self.meshs) <= idx:
            if idx > 2:
                raise ValueError(f"mesh dim with id{idx} is not defined")
            raise ValueError(f"{self._dim_names[idx]} dim is not defined")
        if len(self.meshs[idx]) == 0:
            start, end, n = self.mesh_info[idx]
            self.meshs[idx] = (end - start) * torch.arange(0, n, device=self.device, dtype=self.dtype) / n
        return self.meshs[idx]

    @property
    def x(self):
        return self[0]

    @property
    def y(self):
        return self[1]

    @property
    def z(self):
        return self[2]

    def mesh_grid(self, numpy=False):
        axes = [self[i] for i in range(len(self))]
        if numpy:
            grid = np.meshgrid(*[a.cpu().numpy() for a in axes], indexing="ij")
        else:
            grid = torch.meshgrid(*axes, indexing="ij")
        if len(grid) == 1:
            return grid[0]
        return grid

    def bc_mesh_grid(self, batch_size: int = 1, n_channels: int = 1, numpy=False):
        grid = self.mesh_grid()
        if not isinstance(grid, (list, tuple)):
            grid = [grid]
        out = [g.reshape(1, 1, *g.shape).repeat(batch_size, n_channels, *[1] * g.dim()) for g in grid]
        if numpy:
            out = [o.cpu().numpy() for o in out]
        if len(out) == 1:
            return out[0]
        return out

    def to(self, device=None, dtype=None):
        self.__init__(self.mesh_info, device=device, dtype=dtype)


class FFTFrequency:
    """fftfreq(n, d=(end-start)/n) per axis (mesh.py:148-216)."""

    def __init__(self, mesh_info=(), device=None, dtype=None) -> None:
        self._dim_names = ["x", "y", "z"]
        self.mesh_info = mesh_info
        self.fs = [[] for _ in range(len(mesh_info))]
        self.device, self.dtype = format_device_dtype(device, dtype)

    def __len__(self):
        return len(self.mesh_info)

    def __getitem__(self, idx: int):
        if len(self.fs) <= idx:
            if idx > 2:
                raise ValueError(f"fft frequency with id{idx} is not defined")
            raise ValueError(f"{self._dim_names[idx]} fft frequency is not defined")
        if len(self.fs[idx]) == 0:
            start, end, n = self.mesh_info[idx]
            self.fs[idx] = torch.fft.fftfreq(n, (end - start) / n, device=self.device, dtype=self.dtype)
        return self.fs[idx]

    @property
    def f_x(self):
        return self[0]

    @property
    def f_y(self):
        return self[1]

    @property
    def f_z(self):
        return self[2]

    def to(self, device=None, dtype=None):
        self.__init__(self.mesh_info, device=device, dtype=dtype)


class BroadcastedFFTFrequency:
    """Frequencies reshaped to (1, 1, .., n_i, ..) (mesh.py:219-291)."""

    def __init__(self, fft_frequency: FFTFrequency) -> None:
        self.fft_frequency = fft_frequency
        self.bdks = [[] for _ in range(len(fft_frequency))]
        self._dim_names = fft_frequency._dim_names
        self.mesh_shape = tuple(i[-1] for i in fft_frequency.mesh_info)
        self._bf_vector = None

    def __len__(self):
        return len(self.fft_frequency)

    def __getitem__(self, idx: int):
        if len(self) <= idx:
            if idx > 2:
                raise ValueError(f"fft frequency with id{idx} is not defined")
            raise ValueError(f"{self._dim_names[idx]} fft frequency is not defined")
        if len(self.bdks[idx]) == 0:
            shape = [1] * (len(self.fft_frequency) + 2)
            shape[idx + 2] = self.fft_frequency[idx].shape[0]
            self.bdks[idx] = self.fft_frequency[idx].reshape(*shape)
        return self.bdks[idx]

    @property
    def bf_vector(self):
        if self._bf_vector is None:
            full = (1, 1) + self.mesh_shape
            parts = []
            for i in range(len(self)):
                rep = list(full)
                rep[i + 2] = 1
                parts.append(self[i].repeat(*rep))
            self._bf_vector = torch.cat(parts, dim=1)
        return self._bf_vector

    @property
    def bf_x(self):
        return self[0]

    @property
    def bf_y(self):
        return self[1]

    @property
    def bf_z(self):
        return self[2]

    def to(self, device=None, dtype=None):
        self.fft_frequency.to(device=device, dtype=dtype)
        self.__init__(self.fft_frequency)


class FourierMesh:
    """Frequencies, derivative symbols, dealias mask and transforms of a periodic box
    (mesh.py:294-504).  fft/ifft run through libtfsm on the GPU."""

    def __init__(self, mesh: Union[Sequence[Tuple[float, float, int]], MeshGrid], device=None, dtype=None) -> None:
        if isinstance(mesh, MeshGrid):
            self.mesh_info = mesh.mesh_info
            self.device = default(device, mesh.device)
            self.dtype = default(dtype, mesh.dtype)
            self.device, self.dtype = format_device_dtype(self.device, self.dtype)
        else:
            self.mesh_info = mesh
            self.device, self.dtype = format_device_dtype(device, dtype)
        self.f = FFTFrequency(self.mesh_info, device=self.device, dtype=self.dtype)
        self.bf = BroadcastedFFTFrequency(self.f)
        self.n_dim = len(self.mesh_info)
        self.fft_dim = tuple(-1 * (i + 1) for i in range(self.n_dim))
        self._default_rel_freq_threshold = 2 / 3

    def set_default_rel_freq_threshold(self, threshold: float):
        self._default_rel_freq_threshold = threshold
        self.low_pass_filter.cache_clear()

    @property
    def f_x(self):
        return self.f.f_x

    @property
    def f_y(self):
        return self.f.f_y

    @property
    def f_z(self):
        return self.f.f_z

    @property
    def bf_x(self):
        return self.bf.bf_x

    @property
    def bf_y(self):
        return self.bf.bf_y

    @property
    def bf_z(self):
        return self.bf.bf_z

    @property
    def bf_vector(self):
        return self.bf.bf_vector

    @lru_cache()
    def grad(self, dim_i: int, order: int):
        return (2j * torch.pi * self.bf[dim_i]) ** order

    @lru_cache()
    def laplacian(self):
        return self.nabla(2)

    @lru_cache()
    def invert_laplacian(self):
        lap = self.laplacian()
        return torch.where(lap == 0, 1.0, 1 / lap)

    @lru_cache()
    def nabla(self, order: int = 1):
        return sum([self.grad(dim_i, order) for dim_i in range(len(self.bf))])

    @lru_cache()
    def invert_nabla(self, order: int = 1):
        nab = self.nabla(order)
        return torch.where(nab == 0, 1.0, 1 / nab)

    @lru_cache()
    def nabla_vector(self, order: int):
        return (2j * torch.pi * self.bf.bf_vector) ** order

    @lru_cache()
    def low_pass_filter(self, rel_freq_threshold: Optional[float] = None):
        rel = default(rel_freq_threshold, self._default_rel_freq_threshold)
        mask = torch.ones_like(self.nabla().real)
        for i in range(len(self.bf)):
            abs_f = self.bf[i].abs()
            mask *= torch.where(abs_f > abs_f.max() * rel, 0, 1)
        return mask.to(device=self.device, dtype=self.dtype)

    @lru_cache()
    def abs_low_pass_filter(self, abs_freq_threshold: int):
        mask = torch.ones_like(self.nabla().real)
        for i in range(len(self.bf)):
            abs_f = self.bf[i].abs()
            mask *= torch.where(abs_f > abs_freq_threshold, 0, 1)
        return mask.to(device=self.device, dtype=self.dtype)

    # -------------------------------------------------------------- transforms on the engine
    def _plan_for(self, u: torch.Tensor):
        from .engine import require_gpu_tensor, transform_plan

        require_gpu_tensor(u, "FourierMesh transform input")
        if u.dtype in (torch.float64, torch.complex128):
            raise TypeError("the MI355X engine computes in float32/complex64; got " + str(u.dtype))
        shape = tuple(u.shape[-self.n_dim:])
        if shape != tuple(int(m[2]) for m in self.mesh_info):
            raise ValueError(f"value shape {tuple(u.shape)} does not match the mesh")
        return transform_plan(self.mesh_info, u.device)

    def fft(self, u):
        """torch.fft.fftn over the last n_dim dims, unnormalised (mesh.py:481-485)."""
        plan = self._plan_for(u)
        lead = u.shape[: u.dim() - self.n_dim]
        nb = int(np.prod(lead)) if len(lead) else 1
        if u.is_complex():
            x = u.reshape((nb, 1) + tuple(u.shape[-self.n_dim:]))
            return plan.c2c(x, -1).reshape(u.shape)
        x = u.reshape((nb, 1) + tuple(u.shape[-self.n_dim:]))
        half = plan.forward(x)
        return plan.half_to_full(half).reshape(u.shape)

    def ifft(self, u_fft):
        """torch.fft.ifftn over the last n_dim dims, scaled 1/prod(N) (mesh.py:487-491)."""
        plan = self._plan_for(u_fft)
        lead = u_fft.shape[: u_fft.dim() - self.n_dim]
        nb = int(np.prod(lead)) if len(lead) else 1
        x = u_fft.to(torch.complex64).reshape((nb, 1) + tuple(u_fft.shape[-self.n_dim:]))
        return plan.c2c(x, +1).reshape(u_fft.shape)

    def to(self, device=None, dtype=None):
        self.__init__(self.mesh_info, device=device, dtype=dtype)
        for fn in (self.grad, self.laplacian, self.invert_laplacian, self.nabla, self.invert_nabla,
                   self.nabla_vector, self.low_pass_filter):
            fn.cache_clear()


def mesh_shape(mesh, batch_size: int = 1, n_channel: int = 1) -> Tuple:
    """(batch_size, n_channel, n_1, ..., n_D) (mesh.py:507-527)."""
    if isinstance(mesh, (FourierMesh, MeshGrid)):
        mesh = mesh.mesh_info
    return tuple([batch_size, n_channel] + [m[2] for m in mesh])
