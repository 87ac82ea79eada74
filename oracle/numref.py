"""numref: independent NumPy float64 restatement of the ETD stepping path (oracle; test infrastructure).

It shares no code with the reference or with refpath.  Differences in formulation, chosen so that a
dropped term, a wrong sign or a transposed index shows up as a mismatch against refpath-float64:
  * half spectrum (numpy.fft.rfftn / irfftn over the last D axes) instead of the full complex one;
  * derivative symbols built from integer wavenumbers k (2 pi k / length), not from fftfreq;
  * the dealias mask from the integer rule |k_i| <= floor(rate * N_i / 2) per axis
    (torchfsm mesh.py:443-461 keeps |f| <= rate * max|f|, the same set for the sizes used);
  * phi-function coefficients from closed forms with Taylor series near z = 0, not contour means
    (the contour mean of _uncached.py:34-43 converges to these for every z).
"""

import math

import numpy as np


def axis_wavenumbers(n, length):
    """2 pi k / length for k = 0..n/2-1, -n/2..-1 (numpy.fft.fftfreq order)."""
    k = np.fft.fftfreq(n, 1.0 / n)
    return 2 * math.pi * k / length


def half_shape(shape):
    return tuple(shape[:-1]) + (shape[-1] // 2 + 1,)


class Box:
    """Periodic box of D dims; arrays are [B, C, N_1, ..., N_D] (spatial) or half spectra."""

    def __init__(self, mesh_info):
        self.mesh_info = [tuple(m) for m in mesh_info]
        self.D = len(self.mesh_info)
        self.N = tuple(int(m[2]) for m in self.mesh_info)
        self.length = tuple(float(m[1] - m[0]) for m in self.mesh_info)
        self.axes = tuple(range(-self.D, 0))
        ks = []
        for d in range(self.D):
            k = axis_wavenumbers(self.N[d], self.length[d])
            if d == self.D - 1:
                k = k[: self.N[d] // 2 + 1].copy()
                if self.N[d] % 2 == 0:
                    k[-1] = -k[-1]  # Nyquist carries the negative wavenumber (fftfreq)
            shape = [1] * self.D
            shape[d] = k.shape[0]
            ks.append(k.reshape(shape))
        self.k = ks  # broadcastable over the half spectrum

    def rfft(self, u):
        return np.fft.rfftn(u, axes=self.axes)

    def irfft(self, uh):
        return np.fft.irfftn(uh, s=self.N, axes=self.axes)

    def lap(self):
        return -sum(k**2 for k in self.k)

    def mask(self, rate=2 / 3):
        m = np.ones(half_shape(self.N))
        for d in range(self.D):
            n = self.N[d]
            kint = np.fft.fftfreq(n, 1.0 / n)
            if d == self.D - 1:
                kint = kint[: n // 2 + 1]
            keep = np.abs(kint) <= math.floor(rate * (n // 2) + 1e-9)
            shape = [1] * self.D
            shape[d] = keep.shape[0]
            m = m * keep.reshape(shape)
        return m


def vorticity_convection(box: Box, wh):
    """(u.grad)w with u = (d_y psi, -d_x psi), psi = -lap^-1 w (lap^-1(0) := 1)."""
    lap = box.lap()
    inv = np.where(lap == 0, 1.0, 1.0 / np.where(lap == 0, 1.0, lap))
    psi = -wh * inv
    kx, ky = box.k
    ux = box.irfft(1j * ky * psi)
    uy = box.irfft(-1j * kx * psi)
    wx = box.irfft(1j * kx * wh)
    wy = box.irfft(1j * ky * wh)
    return box.rfft(ux * wx + uy * wy)


def convection(box: Box, uh):
    """Sum_i u_i d_i u_c for a C = D vector field given as half spectra [B, C, ...]."""
    u = box.irfft(uh)
    out = np.zeros_like(uh)
    for c in range(uh.shape[1]):
        acc = 0.0
        for i in range(box.D):
            acc = acc + u[:, i] * box.irfft(1j * box.k[i] * uh[:, c])
        out[:, c] = box.rfft(acc)
    return out


def phi(z, order):
    """phi_order(z) = sum_{m>=0} z^m / (m + order)!, closed form away from 0, Taylor near 0."""
    z = np.asarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    small = np.abs(z) < 0.5
    zs = z[small]
    acc = np.zeros_like(zs)
    term = np.ones_like(zs) / math.factorial(order)
    for m in range(30):
        acc = acc + term
        term = term * zs / (m + 1 + order)
    out[small] = acc
    zb = z[~small]
    val = np.exp(zb)
    for j in range(order):
        val = (val - 1.0 / math.factorial(j)) / zb
    out[~small] = val
    return out


def coefficients(scheme, dt, L):
    """Per-mode coefficients of the ETD schemes, written with phi functions of z = dt L."""
    z = dt * L.astype(np.complex128)
    c = {"E": np.exp(z)}
    if scheme in ("ETDRK1", "SETDRK1"):
        c["c1"] = dt * phi(z, 1)
    elif scheme in ("ETDRK2", "SETDRK2"):
        c["c1"] = dt * phi(z, 1)
        c["c2"] = dt * phi(z, 2)
    elif scheme == "SETDRK3":
        c["E2"] = np.exp(z / 2)
        c["c1"] = dt / 2 * phi(z / 2, 1)
        c["c2"] = dt * phi(z, 1)
        c["c3"] = dt * (phi(z, 1) - 3 * phi(z, 2) + 4 * phi(z, 3))
        c["c4"] = dt * 4 * (phi(z, 2) - 2 * phi(z, 3))
        c["c5"] = dt * (-phi(z, 2) + 4 * phi(z, 3))
    elif scheme == "SETDRK4":
        c["E2"] = np.exp(z / 2)
        c["c1"] = dt / 2 * phi(z / 2, 1)
        c["c4"] = dt * (phi(z, 1) - 3 * phi(z, 2) + 4 * phi(z, 3))
        c["c5"] = dt * (phi(z, 2) - 2 * phi(z, 3))
        c["c6"] = dt * (-phi(z, 2) + 4 * phi(z, 3))
    elif scheme != "ETDRK0":
        raise ValueError(scheme)
    return {k: v.real for k, v in c.items()}


class Problem:
    """L = sum of linear symbols; N = sum of nonlinear terms on the dealiased state + sources."""

    def __init__(self, mesh_info, terms, n_channel, rate=2 / 3):
        self.box = Box(mesh_info)
        self.C = n_channel
        lap = self.box.lap()
        L = np.zeros(half_shape(self.box.N))
        self.nl = []
        for coef, kind, payload in terms:
            if kind == "lap":
                L = L + coef * lap
            elif kind == "biharm":
                L = L + coef * lap * lap
            elif kind == "unit":
                L = L + coef
            elif kind in ("conv", "vc"):
                self.nl.append((coef, kind, None))
            elif kind == "src":
                self.nl.append((coef, kind, self.box.rfft(np.asarray(payload, dtype=np.float64))))
            else:
                raise ValueError(kind)
        self.L = np.broadcast_to(L, (1, n_channel) + L.shape)
        self.mask = self.box.mask(rate)
        self.has_nl = bool(self.nl)

    def N(self, uh):
        out = np.zeros_like(uh)
        for coef, kind, payload in self.nl:
            if kind == "vc":
                out = out + coef * vorticity_convection(self.box, uh * self.mask)
            elif kind == "conv":
                out = out + coef * convection(self.box, uh * self.mask)
            else:
                out = out + coef * payload
        return out


def step(prob: Problem, scheme, c, uh):
    N = prob.N
    if scheme == "ETDRK0":
        return c["E"] * uh
    if scheme in ("ETDRK1", "SETDRK1"):
        return c["E"] * uh + c["c1"] * N(uh)
    if scheme in ("ETDRK2", "SETDRK2"):
        n0 = N(uh)
        a = c["E"] * uh + c["c1"] * n0
        return a + c["c2"] * (N(a) - n0)
    if scheme == "SETDRK3":
        n0 = N(uh)
        a = c["E2"] * uh + c["c1"] * n0
        na = N(a)
        b = c["E"] * uh + c["c2"] * (2 * na - n0)
        nb = N(b)
        return c["E"] * uh + c["c3"] * n0 + c["c4"] * na + c["c5"] * nb
    if scheme == "SETDRK4":
        n0 = N(uh)
        a = c["E2"] * uh + c["c1"] * n0
        na = N(a)
        b = c["E2"] * uh + c["c1"] * na
        nb = N(b)
        cc = c["E2"] * a + c["c1"] * (2 * nb - n0)
        nc = N(cc)
        return c["E"] * uh + c["c4"] * n0 + 2 * c["c5"] * (na + nb) + c["c6"] * nc
    raise ValueError(scheme)


def integrate(mesh_info, terms, u0, dt, steps, scheme="auto", rate=2 / 3, record_every=None):
    """Integrate a real field u0 [B, C, N...] in float64; returns the final field or a trajectory."""
    u0 = np.asarray(u0, dtype=np.float64)
    prob = Problem(mesh_info, terms, u0.shape[1], rate)
    if scheme == "auto":
        scheme = "SETDRK4" if prob.has_nl else "ETDRK0"
    c = coefficients(scheme, dt, prob.L)
    uh = prob.box.rfft(u0)
    frames = []
    for i in range(steps):
        if record_every and i % record_every == 0:
            frames.append(prob.box.irfft(uh))
        uh = step(prob, scheme, c, uh)
    if record_every:
        frames.append(prob.box.irfft(uh))
        return np.stack(frames, axis=1)
    return prob.box.irfft(uh)


def dft_brute(x, sign=-1):
    """O(N^2) DFT along the last axis in float64 (a check for any FFT of small length)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    j = np.arange(n)
    w = np.exp(sign * 2j * math.pi * np.outer(j, j) / n)
    return x @ w.T
