"""Mesh sizes with factors 3, 5 and 7 (mixed-radix Stockham transforms): the FFT passes against
numpy.fft, and integration against the oracle on the same inputs (rel-L2 <= 1e-5 in float32),
including the Kuramoto-Sivashinsky configuration of the reference's gallery (N = 100, L = 60)."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
TOL = 1e-5
MIXED_1D = [48, 96, 100, 112, 192, 200, 384, 400, 768, 800]
MIXED_2D = [48, 96, 100, 112, 192, 200]


def _rel(a, b):
    a = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a).astype(np.complex128)
    b = np.asarray(b.detach().cpu() if isinstance(b, torch.Tensor) else b).astype(np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", MIXED_1D)
def test_r2c_c2r_1d_mixed(dev, n):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n)
    u = rng.standard_normal((3, 1, n)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n)], 3, 1, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    ref = np.fft.rfft(u.astype(np.float64), axis=-1)
    e1 = _rel(uh[..., : n // 2 + 1], ref)
    e2 = _rel(plan.inverse(uh), u)
    e3 = _rel(plan.half_to_full(uh), np.fft.fft(u.astype(np.float64), axis=-1))
    print(f"1D {n}: r2c {e1:.2e} c2r {e2:.2e} full {e3:.2e}")
    assert e1 < 2e-6 and e2 < 2e-6 and e3 < 2e-6


@pytest.mark.parametrize("n0,n1", [(48, 48), (96, 96), (100, 100), (112, 112), (192, 192), (200, 200), (100, 64),
                                   (64, 96)])
def test_r2c_c2r_2d_mixed(dev, n0, n1):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n0 * 7 + n1)
    u = rng.standard_normal((2, 2, n0, n1)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n0), (0, 2, n1)], 2, 2, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    e1 = _rel(uh[..., : n1 // 2 + 1], np.fft.rfftn(u.astype(np.float64), axes=(-2, -1)))
    e2 = _rel(plan.inverse(uh), u)
    print(f"2D {n0}x{n1}: r2c {e1:.2e} c2r {e2:.2e}")
    assert e1 < 3e-6 and e2 < 3e-6


@pytest.mark.parametrize("n", [48, 96, 100])
def test_r2c_c2r_3d_mixed(dev, n):
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n)
    u = rng.standard_normal((1, 1, n, n, n)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n)] * 3, 1, 1, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    e1 = _rel(uh[..., : n // 2 + 1], np.fft.rfftn(u.astype(np.float64), axes=(-3, -2, -1)))
    e2 = _rel(plan.inverse(uh), u)
    print(f"3D {n}^3: r2c {e1:.2e} c2r {e2:.2e}")
    assert e1 < 3e-6 and e2 < 3e-6


def test_ks_gallery_n100(dev):
    """-Laplacian - Biharmonic - Convection at N = 100, L = 60, dt = 0.5, SETDRK4 (the gallery's
    Kuramoto-Sivashinsky run), 20 steps with a recorder, B = 4 random initial states."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.traj_recorder import AutoRecorder

    torch.manual_seed(0)
    u0 = torch.randn(4, 1, 100)
    mi = [(0, 60.0, 100)]
    op = pde.KuramotoSivashinsky()
    rec = AutoRecorder()
    traj = op.integrate(u0.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=0.5, step=20, trajectory_recorder=rec)
    ref = refpath.integrate(refpath.ks_terms(), mi, u0, 0.5, 20, scheme="SETDRK4", record=lambda s: True)
    e = _rel(traj, ref)
    print(f"KS N=100 L=60, 20 steps: rel-L2 {e:.2e}")
    assert traj.shape == ref.shape and e <= TOL


@pytest.mark.parametrize("n", MIXED_1D)
def test_burgers_1d_mixed(dev, n):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.workloads import fourier_modes_field

    u = 0.5 * fourier_modes_field(n, 1, 1, 3, seed=n)
    mi = [(0, 1, n)]
    out = pde.Burgers(0.01).integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=1e-3, step=20)
    ref = refpath.integrate(refpath.burgers_terms(0.01), mi, u, 1e-3, 20, scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"Burgers 1D {n}: rel-L2 {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("n", MIXED_2D)
def test_ns_and_burgers_2d_mixed(dev, n):
    """Vorticity-form NS (Kolmogorov forcing, drag) and 2D Burgers at a mixed-radix size, every
    column tile of the plan."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import field, pde
    from torchfsm_amd.workloads import fourier_modes_field, ns_initial_vorticity

    mesh = tfa.MeshGrid([(0, 2 * math.pi, n)] * 2, device=dev)
    y = mesh.bc_mesh_grid()[1]
    mi = [(0, 2 * math.pi, n)] * 2
    w0 = ns_initial_vorticity(n, 2, seed=n, kmax=n / 4)
    ref = refpath.integrate(refpath.ns_terms(1000, y.cpu()), mi, w0, 1e-3, 10, scheme="SETDRK4")
    for tile in (8, 16):
        op = pde.NavierStokesVorticity(1000, field.kolm_force(y))
        op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=1)
        plan = op._state_dict["integrator"]
        plan.set_tuning(col_tile=tile)
        out = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=10)
        e = _rel(out, ref)
        print(f"NS {n}^2 tile {tile}: rel-L2 {e:.2e}")
        assert e <= TOL
    u = 0.5 * fourier_modes_field(n, 2, 2, 2, seed=n + 1)
    mi1 = [(0, 1, n)] * 2
    out = pde.Burgers(0.01).integrate(u.to(dev), mesh=tfa.MeshGrid(mi1, device=dev), dt=5e-4, step=10)
    ref = refpath.integrate(refpath.burgers_terms(0.01), mi1, u, 5e-4, 10, scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"Burgers 2D {n}^2: rel-L2 {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("n", [48, 96, 100])
def test_burgers_3d_mixed(dev, n):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.workloads import fourier_modes_field

    u = 0.5 * fourier_modes_field(n, 3, 3, 1, seed=n + 3)
    mi = [(0, 1, n)] * 3
    out = pde.Burgers(0.01).integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=5e-4, step=3)
    ref = refpath.integrate(refpath.burgers_terms(0.01), mi, u, 5e-4, 3, scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"Burgers 3D {n}^3: rel-L2 {e:.2e}")
    assert e <= TOL
