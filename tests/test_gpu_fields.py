"""Initial-condition generators (field.py:8-125) on the engine against the oracle restatement fed
the same random draws (same seed, same device).  Tolerance: rel-L2 <= 1e-5."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a.cpu(), dtype=np.float64)
    b = np.asarray(b.cpu(), dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("mi", [[(0, 1, 256)], [(0, 2 * math.pi, 64)] * 2, [(0, 1, 32)] * 3])
def test_diffused_noise(dev, mi):
    from oracle import refpath
    from torchfsm_amd import field

    shape = (3, 2) + tuple(m[2] for m in mi)
    torch.manual_seed(11)
    ours = field.diffused_noise(mi, diffusion_coef=0.01, device=dev, batch_size=3, n_channel=2)
    torch.manual_seed(11)
    u0 = torch.randn(*shape, device=dev)
    ref = refpath.diffused_noise_from(u0, mi, 0.01)
    e = _rel(ours, ref)
    print(f"diffused_noise {shape}: rel-L2 {e:.2e}")
    assert tuple(ours.shape) == shape and e <= 1e-5
    torch.manual_seed(11)
    v = field.diffused_noise(mi, diffusion_coef=0.01, device=dev, batch_size=3, n_channel=2,
                             unit_magnitude=False, unit_variance=True)
    assert torch.allclose(v.flatten(1).std(1), torch.ones(3, device=dev), atol=1e-5)


@pytest.mark.parametrize("mi", [[(0, 1, 128)], [(0, 2 * math.pi, 64)] * 2])
def test_truncated_fourier_series(dev, mi):
    from oracle import refpath
    from torchfsm_amd import field

    shape = (2, 1) + tuple(m[2] for m in mi)
    torch.manual_seed(12)
    ours = field.truncated_fourier_series(mi, freq_threshold=5, device=dev, batch_size=2)
    torch.manual_seed(12)
    mag = torch.rand(*shape, device=dev) * 2.0 + -1.0
    ang = torch.rand(*shape, device=dev) * (2.0 * torch.pi - 0.0) + 0.0
    ref = refpath.truncated_fourier_series_from(mag, ang, mi, 5)
    e = _rel(ours, ref)
    print(f"truncated_fourier_series {shape}: rel-L2 {e:.2e}")
    assert e <= 1e-5
    with pytest.raises(ValueError):
        field.truncated_fourier_series(mi, device=dev, unit_variance=True)
