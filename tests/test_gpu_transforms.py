"""FFT passes of the engine (FourierMesh.fft/.ifft, mesh.py:481-491) against numpy.fft."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_r2c_c2r_1d(dev, n):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n)
    u = rng.standard_normal((3, 1, n)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n)], 3, 1, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    ref = np.fft.rfft(u.astype(np.float64), axis=-1)
    got = uh.cpu().numpy()[..., : n // 2 + 1]
    assert _rel(got, ref) < 2e-6
    back = plan.inverse(uh).cpu().numpy()
    assert _rel(back, u) < 2e-6


@pytest.mark.parametrize("n0,n1", [(16, 16), (32, 32), (64, 64), (128, 128), (256, 256), (512, 512), (64, 128)])
def test_r2c_c2r_2d(dev, n0, n1):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n0 + n1)
    u = rng.standard_normal((2, 2, n0, n1)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n0), (0, 2, n1)], 2, 2, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    ref = np.fft.rfftn(u.astype(np.float64), axes=(-2, -1))
    got = uh.cpu().numpy()[..., : n1 // 2 + 1]
    assert _rel(got, ref) < 3e-6
    back = plan.inverse(uh).cpu().numpy()
    assert _rel(back, u) < 3e-6
    full = plan.half_to_full(uh).cpu().numpy()
    assert _rel(full, np.fft.fftn(u.astype(np.float64), axes=(-2, -1))) < 3e-6
    half2 = plan.full_to_half(plan.half_to_full(uh))
    assert _rel(half2[..., : n1 // 2 + 1].cpu().numpy(), uh[..., : n1 // 2 + 1].cpu().numpy()) < 1e-6
    # an exactly Hermitian spectrum (numpy rfftn, mirrored) passes through full_to_half unchanged
    hr = np.fft.rfftn(u.astype(np.float64), axes=(-2, -1)).astype(np.complex64)
    fr = np.fft.fftn(u.astype(np.float64), axes=(-2, -1)).astype(np.complex64)
    fr[..., : n1 // 2 + 1] = hr
    i0 = (-np.arange(n0)) % n0
    for k in range(n1 // 2 + 1, n1):
        fr[..., :, k] = np.conj(hr[..., i0, n1 - k])
    got = plan.full_to_half(torch.from_numpy(fr).to(dev)).cpu().numpy()[..., : n1 // 2 + 1]
    assert np.array_equal(got, hr)


@pytest.mark.parametrize("n", [(16, 16, 16), (32, 32, 32), (64, 64, 64), (128, 128, 128), (16, 32, 64), (128, 64, 32)])
def test_r2c_c2r_3d(dev, n):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(sum(n))
    u = rng.standard_normal((2, 3) + n).astype(np.float32)
    plan = SpectralPlan([(0, 1, n[0]), (0, 2, n[1]), (0, 3, n[2])], 2, 3, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    axes = (-3, -2, -1)
    ref = np.fft.rfftn(u.astype(np.float64), axes=axes)
    got = uh.cpu().numpy()[..., : n[2] // 2 + 1]
    assert _rel(got, ref) < 3e-6
    back = plan.inverse(uh).cpu().numpy()
    assert _rel(back, u) < 3e-6
    full = plan.half_to_full(uh).cpu().numpy()
    assert _rel(full, np.fft.fftn(u.astype(np.float64), axes=axes)) < 3e-6
    z = (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)).astype(np.complex64)
    zt = torch.from_numpy(z).to(dev)
    assert _rel(plan.c2c(zt, -1).cpu().numpy(), np.fft.fftn(z.astype(np.complex128), axes=axes)) < 3e-6
    assert _rel(plan.c2c(zt, +1).cpu().numpy(), np.fft.ifftn(z.astype(np.complex128), axes=axes)) < 3e-6


@pytest.mark.parametrize("shape", [(128,), (64, 64), (512, 512), (32, 32, 32)])
def test_fourier_mesh_fft_ifft(dev, shape):
    import torchfsm_amd as tfa

    rng = np.random.default_rng(7)
    mesh = tfa.FourierMesh([(0, 2 * math.pi, n) for n in shape], device=dev)
    u = rng.standard_normal((2, 1) + shape).astype(np.float32)
    axes = tuple(range(-len(shape), 0))
    uf = mesh.fft(torch.from_numpy(u).to(dev))
    assert uf.dtype == torch.complex64 and tuple(uf.shape) == u.shape
    assert _rel(uf.cpu().numpy(), np.fft.fftn(u.astype(np.float64), axes=axes)) < 3e-6
    z = (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)).astype(np.complex64)
    zt = torch.from_numpy(z).to(dev)
    assert _rel(mesh.fft(zt).cpu().numpy(), np.fft.fftn(z.astype(np.complex128), axes=axes)) < 3e-6
    assert _rel(mesh.ifft(zt).cpu().numpy(), np.fft.ifftn(z.astype(np.complex128), axes=axes)) < 3e-6


def test_fft_known_answers(dev):
    """Impulse -> all ones; single cosine -> two bins of size N/2; brute-force DFT at N=32."""
    from oracle import numref
    from torchfsm_amd.engine import SpectralPlan

    n = 32
    plan = SpectralPlan([(0, 1, n)], 1, 1, device=dev)
    imp = torch.zeros(1, 1, n, device=dev)
    imp[..., 0] = 1
    assert torch.allclose(plan.forward(imp)[..., : n // 2 + 1], torch.ones(n // 2 + 1, dtype=torch.complex64, device=dev))
    x = torch.arange(n, device=dev, dtype=torch.float32)
    c = torch.cos(2 * math.pi * 3 * x / n).reshape(1, 1, n)
    spec = plan.forward(c).cpu().numpy()[0, 0, : n // 2 + 1]
    expect = np.zeros(n // 2 + 1, complex)
    expect[3] = n / 2
    assert np.abs(spec - expect).max() < 1e-4
    rng = np.random.default_rng(3)
    u = rng.standard_normal((1, 1, n)).astype(np.float32)
    got = plan.forward(torch.from_numpy(u).to(dev)).cpu().numpy()[0, 0, : n // 2 + 1]
    assert _rel(got, numref.dft_brute(u[0, 0])[: n // 2 + 1]) < 2e-6
