"""Seeded synthetic inputs of the benchmark and test workloads (generated on the CPU in float64,
cast to float32, identical for every path that consumes them)."""

import math

import torch


def ns_initial_vorticity(n: int, batch: int, seed: int = 0, kmax: float = 32.0, dtype=torch.float32):
    """Gaussian random vorticity on [0, 2 pi)^2: complex normal modes with amplitude 1/|k| for
    1 <= |k| <= kmax (velocity energy spectrum ~ k^-3), zero outside; zero mean, unit RMS per sample."""
    g = torch.Generator().manual_seed(seed)
    k = torch.fft.fftfreq(n, 1.0 / n, dtype=torch.float64)
    kk = torch.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    amp = torch.where((kk >= 1) & (kk <= kmax), 1.0 / torch.clamp(kk, min=1.0), torch.zeros_like(kk))
    re = torch.randn(batch, 1, n, n, generator=g, dtype=torch.float64)
    im = torch.randn(batch, 1, n, n, generator=g, dtype=torch.float64)
    w = torch.fft.ifft2((re + 1j * im) * amp).real
    w = w - w.mean(dim=(-1, -2), keepdim=True)
    w = w / torch.sqrt((w**2).mean(dim=(-1, -2), keepdim=True))
    return w.to(dtype)


def fourier_modes_field(n: int, dims: int, channels: int, batch: int, seed: int = 0, nmodes: int = 8, kabs: int = 4,
                        length: float = 1.0, dtype=torch.float32):
    """Per (batch, channel): sum of nmodes cosines a cos(2 pi k.x / length + phi) with integer k,
    1 <= |k| <= kabs (rejection sampled), a ~ U(-1, 1), phi ~ U(0, 2 pi).  `n` is the points per
    axis, or a tuple of `dims` sizes for a non-cubic box."""
    g = torch.Generator().manual_seed(seed)
    ns = tuple(n) if isinstance(n, (tuple, list)) else (n,) * dims
    if len(ns) != dims:
        raise ValueError("one size per axis expected")
    xs = [torch.arange(nd, dtype=torch.float64) * length / nd for nd in ns]
    grids = torch.meshgrid(*xs, indexing="ij")
    out = torch.zeros(batch, channels, *ns, dtype=torch.float64)
    for b in range(batch):
        for c in range(channels):
            m = 0
            while m < nmodes:
                kvec = torch.randint(-kabs, kabs + 1, (dims,), generator=g)
                kn = float(torch.sqrt((kvec.double() ** 2).sum()))
                if kn < 1 or kn > kabs:
                    continue
                a = float(torch.rand(1, generator=g, dtype=torch.float64) * 2 - 1)
                ph = float(torch.rand(1, generator=g, dtype=torch.float64) * 2 * math.pi)
                arg = sum(2 * math.pi * int(kvec[d]) * grids[d] / length for d in range(dims))
                out[b, c] += a * torch.cos(arg + ph)
                m += 1
    return out.to(dtype)


def burgers1d_initial(n: int = 128, dtype=torch.float32):
    """u0 = sin(2 pi x) + 0.5 on [0, 1) (README example of the reference)."""
    x = torch.arange(n, dtype=dtype) / n
    return (torch.sin(2 * math.pi * x) + 0.5).reshape(1, 1, n)
