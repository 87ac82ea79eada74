"""refpath: PyTorch restatement of torchfsm's ETD stepping schedule (oracle; test infrastructure).

This module reproduces what the reference computes, in the reference's order and dtypes, so that
on the same CPU it matches the reference's float32 output bit for bit (checked against the golden
fixtures in tests/golden).  The HIP engine's results are checked against it, and it is the
torch-GPU and CPU baseline; it is never imported by the product package.

An operator is described by a list of terms in generator order, each ``(coef, kind, payload)``:
  kind "lap"     Laplacian symbol, concatenated over channels   (_laplacian.py:12-15)
  kind "biharm"  lap*lap, concatenated over channels            (_biharmonic.py:13-16)
  kind "unit"    ones_like(bf_x) (ImplicitSource() drag)        (_source.py:9-17)
  kind "deriv"   SpatialDerivative; payload = (dim, order)      (_spatial_derivative.py:7-20)
  kind "conv"    Convection, dealiased input                    (_convection.py:18-46)
  kind "vc"      VorticityConvection, dealiased input           (_navier_stokes.py:27-46)
  kind "nsp"     NSPressureConvection, dealiased input          (_navier_stokes.py:221-262)
  kind "ks"      KSConvection (remove_mean), dealiased input    (_ks_convection.py:18-38)
  kind "cons"    ConservativeConvection, dealiased input        (_conservative_convection.py:17-27)
  kind "src"     ExplicitSource; payload = spatial source field (_base.py:994-1015)
"""

import math
from typing import Callable, List, Optional, Sequence, Tuple

import torch

Term = Tuple[object, str, Optional[torch.Tensor]]
LINEAR_KINDS = ("lap", "biharm", "unit", "deriv")
NONLINEAR_KINDS = ("conv", "vc", "nsp", "nsp_f", "ks", "ks_keep_mean", "cons", "src")


class Mesh:
    """Wavenumbers, symbols, dealias mask and transforms of one periodic box (mesh.py:148-491)."""

    def __init__(self, mesh_info: Sequence[Tuple[float, float, int]], device="cpu", dtype=torch.float32):
        self.mesh_info = [tuple(m) for m in mesh_info]
        self.device = torch.device(device)
        self.dtype = dtype
        self.n_dim = len(self.mesh_info)
        self.fft_dim = tuple(-(i + 1) for i in range(self.n_dim))  # mesh.py:343
        self._bf = {}
        self._cache = {}

    # FFTFrequency.__getitem__ (mesh.py:184-191): fftfreq(n, d=(end-start)/n)
    def freq(self, i):
        start, end, n = self.mesh_info[i]
        return torch.fft.fftfreq(n, (end - start) / n, device=self.device, dtype=self.dtype)

    # BroadcastedFFTFrequency.__getitem__ (mesh.py:250-254)
    def bf(self, i):
        if i not in self._bf:
            shape = [1] * (self.n_dim + 2)
            shape[i + 2] = self.mesh_info[i][2]
            self._bf[i] = self.freq(i).reshape(*shape)
        return self._bf[i]

    # BroadcastedFFTFrequency.bf_vector (mesh.py:256-266)
    def bf_vector(self):
        if "bfv" not in self._cache:
            full = (1, 1) + tuple(m[2] for m in self.mesh_info)
            parts = []
            for i in range(self.n_dim):
                rep = list(full)
                rep[i + 2] = 1
                parts.append(self.bf(i).repeat(*rep))
            self._cache["bfv"] = torch.cat(parts, dim=1)
        return self._cache["bfv"]

    # FourierMesh.grad (mesh.py:399-404)
    def grad(self, i, order):
        key = ("grad", i, order)
        if key not in self._cache:
            self._cache[key] = (2j * torch.pi * self.bf(i)) ** order
        return self._cache[key]

    # FourierMesh.nabla / laplacian (mesh.py:406-426): python sum over dims starting from 0
    def lap(self):
        if "lap" not in self._cache:
            self._cache["lap"] = sum([self.grad(i, 2) for i in range(self.n_dim)])
        return self._cache["lap"]

    # FourierMesh.invert_laplacian (mesh.py:413-419)
    def inv_lap(self):
        if "ilap" not in self._cache:
            lp = self.lap()
            self._cache["ilap"] = torch.where(lp == 0, 1.0, 1 / lp)
        return self._cache["ilap"]

    # FourierMesh.nabla_vector (mesh.py:436-441)
    def nabla_vector(self, order):
        key = ("nv", order)
        if key not in self._cache:
            self._cache[key] = (2j * torch.pi * self.bf_vector()) ** order
        return self._cache[key]

    # FourierMesh.low_pass_filter (mesh.py:443-461): per-dimension box mask
    def mask(self, rate):
        key = ("mask", rate)
        if key not in self._cache:
            m = torch.ones_like(self.lap().real)
            for i in range(self.n_dim):
                a = self.bf(i).abs()
                m *= torch.where(a > a.max() * rate, 0, 1)
            self._cache[key] = m.to(device=self.device, dtype=self.dtype)
        return self._cache[key]

    def fft(self, u):  # mesh.py:481-485
        return torch.fft.fftn(u, dim=self.fft_dim)

    def ifft(self, u_fft):  # mesh.py:487-491
        return torch.fft.ifftn(u_fft, dim=self.fft_dim)


# ----------------------------------------------------------------------------- terms
def linear_symbol(mesh: Mesh, kind: str, n_channel: int, payload=None):
    if kind == "deriv":  # _spatial_derivative.py:18-20: payload = (dim_index, order)
        return mesh.grad(payload[0], payload[1])
    if kind == "lap":  # _laplacian.py:12-15
        return torch.cat([mesh.lap()] * n_channel, dim=1)
    if kind == "biharm":  # _biharmonic.py:13-16
        return torch.cat([mesh.lap() * mesh.lap()] * n_channel, dim=1)
    if kind == "unit":  # _source.py:9-17
        return torch.ones_like(mesh.bf(0))
    raise ValueError(kind)


def convection(u_fft, mesh: Mesh, u=None):
    """(u . grad) u_c (_convection.py:18-46)."""
    if u is None:
        u = mesh.ifft(u_fft).real
    nabla_u = mesh.nabla_vector(1).unsqueeze(2) * u_fft.unsqueeze(1)
    return mesh.fft((u.unsqueeze(2) * mesh.ifft(nabla_u).real).sum(1))


def vorticity_convection(u_fft, mesh: Mesh, u=None):
    """(u . grad) omega through the streamfunction (_navier_stokes.py:27-46)."""
    psi = -u_fft * mesh.inv_lap()
    ux = mesh.ifft(mesh.grad(1, 1) * psi).real
    uy = mesh.ifft(-mesh.grad(0, 1) * psi).real
    wx = mesh.ifft(mesh.grad(0, 1) * u_fft).real
    wy = mesh.ifft(mesh.grad(1, 1) * u_fft).real
    return mesh.fft(ux * wx + uy * wy)


def ns_pressure_convection(u_fft, mesh: Mesh, u=None):
    """_NSPressureConvectionCore without external force (_navier_stokes.py:221-262):
    p = invlap * sum_i(nabla_i * conv_i);  N = nabla * p - conv."""
    conv = convection(u_fft, mesh, u)
    nab = mesh.nabla_vector(1)
    p = mesh.inv_lap() * torch.sum(nab * conv, dim=1, keepdim=True)
    return nab * p - conv


def ns_pressure_convection_forced(u_fft, mesh: Mesh, force, dealias_rate=2 / 3):
    """_NSPressureConvectionCore with an external force (_navier_stokes.py:238-262): the core runs
    without the closure's dealiasing; it takes the force's spectrum from the undealiased input,
    masks the caller's u_fft IN PLACE (the integrator's stage tensor), subtracts the force from
    the convection before the projection and adds it again after:
    N = grad p - (conv - f) + f,  p = invlap div (conv - f)."""
    u_fft *= mesh.mask(dealias_rate)
    u = mesh.ifft(u_fft).real
    conv = convection(u_fft, mesh, u)
    conv -= force
    nab = mesh.nabla_vector(1)
    p = mesh.inv_lap() * torch.sum(nab * conv, dim=1, keepdim=True)
    return nab * p - conv + force


def ks_convection(u_fft, mesh: Mesh, u=None, remove_mean=True):
    """_KSConvectionCore (_ks_convection.py:18-38): 1/2 sum_i (d_i phi)^2, minus its mean over the
    whole tensor (batch included) when remove_mean."""
    grad_u = mesh.ifft(mesh.nabla_vector(1) * u_fft).real
    re = 0.5 * torch.sum(grad_u**2, dim=1, keepdim=True)
    if remove_mean:
        re = re - re.mean()
    return mesh.fft(re)


def source_spectrum(src: torch.Tensor):
    """_ExplicitSourceCore.__init__ (_base.py:1002-1005): fftn over all dims after (B, C)."""
    return torch.fft.fftn(src, dim=[i + 2 for i in range(src.dim() - 2)])


def check_terms(mesh: Mesh, terms: List[Term], n_channel: int):
    for _, kind, _ in terms:
        if kind in ("conv", "nsp", "nsp_f") and mesh.n_dim != n_channel:  # _convection.py:57-64
            raise ValueError("convection operator only works for vector field with the same dimension as mesh")
        if kind == "vc" and (mesh.n_dim != 2 or n_channel != 1):  # _navier_stokes.py:56-59
            raise ValueError("Only vorticity in 2Dmesh is supported")


def build_linear(mesh: Mesh, terms: List[Term], n_channel: int):
    """_build_linear_coefs (_base.py:339-357): sum(coef * symbol) in generator order."""
    lin = [(c, k, p) for c, k, p in terms if k in LINEAR_KINDS]
    if not lin:
        return None
    return sum([c * linear_symbol(mesh, k, n_channel, p) for c, k, p in lin])


def build_closure(mesh: Mesh, terms: List[Term], dealias_rate=2 / 3) -> Optional[Callable]:
    """nonlinear_funcs_all (_base.py:359-406): dealias only the inputs of dealiased terms."""
    nl = []
    for c, k, p in terms:
        if k in ("src", "nsp_f"):  # nsp_f payload: the force's spatial field (its explicit sources)
            nl.append((c, k, source_spectrum(p).to(mesh.device)))
        elif k in ("conv", "vc", "nsp", "ks", "ks_keep_mean", "cons"):
            nl.append((c, k, None))
    if not nl:
        return None

    def closure(u_fft):
        result = 0.0
        d_fft = None
        d_u = None
        u = None
        for c, k, p in nl:
            if k in ("conv", "vc", "nsp", "ks", "ks_keep_mean", "cons"):
                if d_fft is None:
                    d_fft = u_fft * mesh.mask(dealias_rate)
                    d_u = mesh.ifft(d_fft).real
                if k == "cons":
                    uu_fft = mesh.fft(d_u.unsqueeze(2) * d_u.unsqueeze(1))
                    result += c * (mesh.nabla_vector(1).unsqueeze(2) * uu_fft).sum(1)
                elif k in ("ks", "ks_keep_mean"):
                    result += c * ks_convection(d_fft, mesh, d_u, remove_mean=(k == "ks"))
                else:
                    f = {"conv": convection, "vc": vorticity_convection, "nsp": ns_pressure_convection}[k]
                    result += c * f(d_fft, mesh, d_u)
            elif k == "nsp_f":  # dealiasing switched off: the raw stage tensor, masked in place
                if u is None:
                    u = mesh.ifft(u_fft).real
                result += c * ns_pressure_convection_forced(u_fft, mesh, p, dealias_rate)
            else:  # explicit source: no dealiasing, u computed but unused (_base.py:394-401)
                if u is None:
                    u = mesh.ifft(u_fft).real
                result += c * p
        return result

    return closure


# ----------------------------------------------------------------------------- integrators
def roots_of_unity(m, device, dtype):
    """_uncached.py:8-15: exp(2 pi i (j - 1/2)/M), j = 1..M."""
    return torch.exp(2j * torch.pi * (torch.arange(1, m + 1, device=device, dtype=dtype) - 0.5) / m)


class Scheme:
    """Coefficients and step of one ETD scheme, following the reference integrator classes."""

    def __init__(self, name: str, dt: float, L, closure, n_points=16, radius=1.0):
        self.name = name
        self.dt = dt
        self.N = closure
        self.E = torch.exp(dt * L)  # _etdrk.py:21, _uncached.py:30
        if name == "ETDRK0":
            return
        if name in ("ETDRK1", "ETDRK2"):  # _etdrk.py:43-45, 66-70
            self.c1 = torch.where(L == 0, dt, (self.E - 1) / L)
            if name == "ETDRK2":
                self.c2 = torch.where(L == 0, dt / 2, (self.E - 1 - L * dt) / (L**2 * dt))
            return
        lr = radius * roots_of_unity(n_points, L.device, L.real.dtype) + L.unsqueeze(-1) * dt  # _uncached.py:34-43

        def mean(z):
            return dt * torch.mean(z, axis=-1).real

        if name == "SETDRK1":  # _uncached.py:71-72
            self.c1 = mean((torch.exp(lr) - 1) / lr)
        elif name == "SETDRK2":  # _uncached.py:102-104
            self.c1 = mean((torch.exp(lr) - 1) / lr)
            self.c2 = mean((torch.exp(lr) - 1 - lr) / lr**2)
        elif name == "SETDRK3":  # _uncached.py:135-156
            self.E2 = torch.exp(0.5 * dt * L)
            self.c1 = mean((torch.exp(lr / 2) - 1) / lr)
            self.c2 = mean((torch.exp(lr) - 1) / lr)
            self.c3 = mean((-4 - lr + torch.exp(lr) * (4 - 3 * lr + lr**2)) / (lr**3))
            self.c4 = mean((4.0 * (2.0 + lr + torch.exp(lr) * (-2 + lr))) / (lr**3))
            self.c5 = mean((-4 - 3 * lr - lr**2 + torch.exp(lr) * (4 - lr)) / (lr**3))
        elif name == "SETDRK4":  # _uncached.py:191-211
            self.E2 = torch.exp(0.5 * dt * L)
            self.c1 = mean((torch.exp(lr / 2) - 1) / lr)
            self.c2 = self.c1
            self.c3 = self.c1
            self.c4 = mean((-4 - lr + torch.exp(lr) * (4 - 3 * lr + lr**2)) / (lr**3))
            self.c5 = mean((2 + lr + torch.exp(lr) * (-2 + lr)) / (lr**3))
            self.c6 = mean((-4 - 3 * lr - lr**2 + torch.exp(lr) * (4 - lr)) / (lr**3))
        else:
            raise ValueError(name)

    def step(self, u):
        n = self.name
        if n == "ETDRK0":  # _etdrk.py:23-27
            return self.E * u
        if n in ("ETDRK1", "SETDRK1"):  # _etdrk.py:47-51, _setdrk_step.py:5-11
            return self.E * u + self.c1 * self.N(u)
        if n in ("ETDRK2", "SETDRK2"):  # _etdrk.py:72-82, _setdrk_step.py:14-25
            n0 = self.N(u)
            a = self.E * u + self.c1 * n0
            na = self.N(a)
            return a + self.c2 * (na - n0)
        if n == "SETDRK3":  # _setdrk_step.py:28-52
            n0 = self.N(u)
            a = self.E2 * u + self.c1 * n0
            na = self.N(a)
            b = self.E * u + self.c2 * (2 * na - n0)
            nb = self.N(b)
            return self.E * u + self.c3 * n0 + self.c4 * na + self.c5 * nb
        # SETDRK4 (_setdrk_step.py:55-82)
        n0 = self.N(u)
        a = self.E2 * u + self.c1 * n0
        na = self.N(a)
        b = self.E2 * u + self.c2 * na
        nb = self.N(b)
        c = self.E2 * a + self.c3 * (2 * nb - n0)
        nc = self.N(c)
        return self.E * u + self.c4 * n0 + self.c5 * 2 * (na + nb) + self.c6 * nc


# Butcher tableaux of the reference's explicit Runge-Kutta family (_rk.py:82-255): each row of
# CA is [c_i, a_i1, ...], B the update weights.
RK_TABLEAUX = {
    "Euler": ([[1.0]], [1.0]),
    "Midpoint": ([[1 / 2, 1 / 2]], [0, 1]),
    "Heun12": ([[1, 1]], [1 / 2, 1 / 2]),
    "Ralston12": ([[2 / 3, 2 / 3]], [1 / 4, 3 / 4]),
    "BogackiShampine23": ([[1 / 2, 1 / 2], [3 / 4, 0, 3 / 4], [1, 2 / 9, 1 / 3, 4 / 9]], [2 / 9, 1 / 3, 4 / 9, 0]),
    "RK4": ([[1 / 2, 1 / 2], [1 / 2, 0, 1 / 2], [1, 0, 0, 1]], [1 / 6, 1 / 3, 1 / 3, 1 / 6]),
    "RK4_38Rule": ([[1 / 3, 1 / 3], [2 / 3, -1 / 3, 1], [1, -1, 1, 1]], [1 / 8, 3 / 8, 3 / 8, 1 / 8]),
    "Dorpi45": ([[1 / 5, 1 / 5], [3 / 10, 3 / 40, 9 / 40], [4 / 5, 44 / 45, -56 / 15, 32 / 9],
                 [8 / 9, 19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
                 [1, 9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
                 [1, 35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
                [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]),
    "Fehlberg45": ([[1 / 4, 1 / 4], [3 / 8, 3 / 32, 9 / 32], [12 / 13, 1932 / 2197, -7200 / 2197, 7296 / 2197],
                    [1, 439 / 216, -8, 3680 / 513, -845 / 4104], [1 / 2, -8 / 27, 2, -3544 / 2565, 1859 / 4104, -11 / 40]],
                   [16 / 135, 0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55]),
    "CashKarp45": ([[1 / 5, 1 / 5], [3 / 10, 3 / 40, 9 / 40], [3 / 5, 3 / 10, -9 / 10, 6 / 5],
                    [1, -11 / 54, 5 / 2, -70 / 27, 35 / 27],
                    [7 / 8, 1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]],
                   [37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771]),
}


def rk_step(f: Callable, x, dt, name: str):
    """One explicit Runge-Kutta step of _rk.py:43-58 (fixed step, every stage evaluated)."""
    ca, b = RK_TABLEAUX[name]
    ks = [f(x)]
    for row in ca:
        ks.append(f(x + dt * sum([a_i * k for a_i, k in zip(row[1:], ks)])))
    return x + dt * sum([b_i * k for b_i, k in zip(b, ks)])


class RKScheme:
    """Explicit Runge-Kutta stepping of the operator f(u) = L u + N(u) (_base.py:408-439)."""

    def __init__(self, name: str, dt: float, L, closure):
        self.name, self.dt = name, dt
        if closure is None:
            self.f = lambda u: L * u
        elif L is None:
            self.f = closure
        else:
            self.f = lambda u: L * u + closure(u)

    def step(self, u):
        return rk_step(self.f, u, self.dt, self.name)


def choose_scheme(terms: List[Term], scheme: str) -> str:
    """_build_integrator 'auto' (_base.py:451-455): ETDRK0 if linear else SETDRK4."""
    if scheme != "auto":
        return scheme
    return "SETDRK4" if any(k in NONLINEAR_KINDS for _, k, _ in terms) else "ETDRK0"


def integrate(terms: List[Term], mesh_info, u0, dt, step, scheme="auto", record=None,
              return_in_fourier=False, dealias_rate=2 / 3, u0_fft=None, n_points=16, radius=1.0):
    """OperatorLike.integrate (_base.py:676-751).

    ``record`` is None (return the final frame) or a predicate step -> bool; recorded frames are
    stacked on dim 1 and inverse transformed like AutoRecorder.trajectory (traj_recorder.py:112-125).
    """
    x = u0 if u0 is not None else u0_fft
    mesh = Mesh(mesh_info, device=x.device, dtype=(u0.dtype if u0 is not None else x.real.dtype))
    n_channel = x.shape[1]
    check_terms(mesh, terms, n_channel)
    L = build_linear(mesh, terms, n_channel)
    closure = build_closure(mesh, terms, dealias_rate)
    name = choose_scheme(terms, scheme)
    if name in RK_TABLEAUX:  # _base.py:492-500
        sch = RKScheme(name, dt, L, closure)
    else:
        if name != "ETDRK0" and L is None:  # _base.py:473-478
            L = torch.tensor([0.0], dtype=mesh.dtype, device=mesh.device)
        sch = Scheme(name, dt, L, closure, n_points, radius)
    u_fft = mesh.fft(u0) if u0_fft is None else u0_fft
    frames = []
    i = -1
    for i in range(step):
        if record is not None and record(i):
            frames.append(u_fft.clone())
        u_fft = sch.step(u_fft)
    if record is not None:
        if record(i + 1):
            frames.append(u_fft.clone())
        if not frames:
            return None
        traj = torch.stack(frames, dim=1)
        if return_in_fourier:
            return traj
        return torch.fft.ifftn(traj, dim=tuple(-(j + 1) for j in range(traj.dim() - 3))).real
    return u_fft if return_in_fourier else mesh.ifft(u_fft).real


def apply(terms: List[Term], mesh_info, u, dealias_rate=2 / 3, return_in_fourier=False):
    """OperatorLike.__call__ (_base.py:753-790) via _build_operator (_base.py:408-439)."""
    mesh = Mesh(mesh_info, device=u.device, dtype=u.dtype)
    n_channel = u.shape[1]
    check_terms(mesh, terms, n_channel)
    L = build_linear(mesh, terms, n_channel)
    closure = build_closure(mesh, terms, dealias_rate)
    u_fft = mesh.fft(u)
    if closure is None:
        out = L * u_fft
    elif L is None:
        out = closure(u_fft)
    else:
        out = L * u_fft + closure(u_fft)
    return out if return_in_fourier else mesh.ifft(out).real


# ----------------------------------------------------------------------------- diagnostic operators
def div(u_fft, mesh: Mesh):
    """_DivCore (_div.py:17-28)."""
    return torch.sum(mesh.nabla_vector(1) * u_fft, dim=1, keepdim=True)


def grad(u_fft, mesh: Mesh):
    """Grad: linear coefficient nabla_vector(1) times u^ (_grad.py:12-15)."""
    return mesh.nabla_vector(1) * u_fft


def curl(u_fft, mesh: Mesh):
    """_Curl2DCore / _Curl3DCore (_curl.py:18-56)."""
    if mesh.n_dim == 2:
        return mesh.grad(0, 1) * u_fft[:, 1:2] - mesh.grad(1, 1) * u_fft[:, 0:1]
    return torch.cat([mesh.grad(1, 1) * u_fft[:, 2:3] - mesh.grad(2, 1) * u_fft[:, 1:2],
                      mesh.grad(2, 1) * u_fft[:, 0:1] - mesh.grad(0, 1) * u_fft[:, 2:3],
                      mesh.grad(0, 1) * u_fft[:, 1:2] - mesh.grad(1, 1) * u_fft[:, 0:1]], dim=1)


def vorticity2velocity_coef(mesh: Mesh):
    """_Vorticity2VelocityCore (_navier_stokes.py:75-91)."""
    n0, n1 = mesh.mesh_info[0][-1], mesh.mesh_info[1][-1]
    return -1 * mesh.inv_lap() * torch.cat([mesh.grad(1, 1).repeat([1, 1, n0, 1]),
                                           -mesh.grad(0, 1).repeat([1, 1, 1, n1])], dim=1)


def pressure_of_convection(conv, mesh: Mesh):
    """p = -1 * div(conv) * invlap (_navier_stokes.py:140-144, 202-206)."""
    p = torch.sum(mesh.nabla_vector(1) * conv, dim=1, keepdim=True)
    return -1 * p * mesh.inv_lap()


def vorticity2pressure(w_fft, mesh: Mesh, dealias_rate=2 / 3):
    """Vorticity2Pressure()(u_fft=w^) (_navier_stokes.py:119-170): closure mask, velocity,
    Convection core, pressure."""
    wd = w_fft * mesh.mask(dealias_rate)
    vel = wd * vorticity2velocity_coef(mesh)
    return pressure_of_convection(convection(vel, mesh, mesh.ifft(vel).real), mesh)


def velocity2pressure(u_fft, mesh: Mesh, dealias_rate=2 / 3):
    """Velocity2Pressure()(u_fft=u^) (_navier_stokes.py:174-217)."""
    ud = u_fft * mesh.mask(dealias_rate)
    return pressure_of_convection(convection(ud, mesh, mesh.ifft(ud).real), mesh)


def conservative_convection(u_fft, mesh: Mesh, dealias_rate=2 / 3):
    """ConservativeConvection()(u_fft=u^) (_conservative_convection.py:17-27), closure mask first."""
    ud = u_fft * mesh.mask(dealias_rate)
    u = mesh.ifft(ud).real
    uu_fft = mesh.fft(u.unsqueeze(2) * u.unsqueeze(1))
    return (mesh.nabla_vector(1).unsqueeze(2) * uu_fft).sum(1)


# ----------------------------------------------------------------------------- initial-condition generators
def normalise_field(u, zero_centered=True, unit_variance=False, unit_magnitude=True):
    """Normalisation tail of diffused_noise / truncated_fourier_series (field.py:55-62, 117-124)."""
    dims = [i for i in range(1, u.ndim)]
    if zero_centered:
        u = u - u.mean(dim=dims, keepdim=True)
    if unit_variance:
        u = u / u.std(dim=dims, keepdim=True)
    if unit_magnitude:
        u = u / torch.max(u.abs().view(u.size(0), -1), dim=1).values.view([u.shape[0]] + [1] * (u.ndim - 1))
    return u


def diffused_noise_from(u0, mesh_info, diffusion_coef=1.0, **norm):
    """diffused_noise (field.py:8-63) given its Gaussian draw u0: one ETDRK0 step (dt = 1) of
    diffusion_coef * Laplacian(), then normalisation."""
    out = integrate([(diffusion_coef, "lap", None)], mesh_info, u0, 1, 1)
    return normalise_field(out, **norm)


def truncated_fourier_series_from(magnitude, angle, mesh_info, freq_threshold=5, **norm):
    """truncated_fourier_series (field.py:65-125) given its uniform draws (already scaled to the
    amplitude and angle ranges)."""
    mesh = Mesh(mesh_info, device=magnitude.device, dtype=magnitude.dtype)
    noise = magnitude * torch.exp(1j * angle)
    mask = torch.ones_like(mesh.lap().real)
    for d in range(mesh.n_dim):  # FourierMesh.abs_low_pass_filter (mesh.py:463-479)
        mask = mask * torch.where(mesh.bf(d).abs() > freq_threshold, 0, 1)
    noise = noise * mask.to(dtype=mesh.dtype)
    return normalise_field(mesh.ifft(noise).real, **norm)


# ----------------------------------------------------------------------------- canonical operators
def burgers_terms(nu) -> List[Term]:
    """pde.Burgers (pde.py:13-25): nu*Laplacian() - Convection()."""
    return [(nu, "lap", None), (-1, "conv", None)]


def ns_velocity_terms(Re) -> List[Term]:
    """pde.NavierStokes without force (pde.py:85-100): NSPressureConvection() + 1/Re Laplacian()."""
    return [(1, "nsp", None), (1 / Re, "lap", None)]


def ks_terms() -> List[Term]:
    """pde.KuramotoSivashinsky (pde.py:27-39): -Laplacian() - Biharmonic() - Convection()."""
    return [(-1, "lap", None), (-1, "biharm", None), (-1, "conv", None)]


def kdv_terms(dispersion_coef=1.0, convection_coef=6.0) -> List[Term]:
    """pde.KortewegDeVries (pde.py:51-64): -c1 SpatialDerivative(0, 3) + c2 Convection()."""
    return [(-dispersion_coef, "deriv", (0, 3)), (convection_coef, "conv", None)]


def ks_highdim_terms(remove_mean=True) -> List[Term]:
    """pde.KuramotoSivashinskyHighDim (pde.py:41-49): -Laplacian() - Biharmonic() - KSConvection()."""
    return [(-1, "lap", None), (-1, "biharm", None), (-1, "ks" if remove_mean else "ks_keep_mean", None)]


def ns_terms(Re, y=None, drag=-0.1, k=4.0, length_scale=1.0) -> List[Term]:
    """pde.NavierStokesVorticity + field.kolm_force (pde.py:66-83, field.py:128-148)."""
    t = [(-1, "vc", None), (1 / Re, "lap", None)]
    if y is not None:
        t += [(drag, "unit", None), (-1, "src", k * torch.cos(k * length_scale * y))]
    return t


def rel_l2(a, b):
    a = torch.as_tensor(a).double()
    b = torch.as_tensor(b).double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
