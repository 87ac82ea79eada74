"""Operators with an imaginary linear symbol (odd-order SpatialDerivative, _spatial_derivative.py)
on the engine against the oracle on the same inputs: the KdV gallery run (pde.KortewegDeVries,
SETDRK4 with the reference's real-part SETDRK coefficients), every ETD scheme and RK4 on KdV, the
linear Airy / advection equations (ETDRK0) in 1D and 2D, Operator.__call__ and
LinearOperator.solve.  Tolerance: rel-L2 <= 1e-5 (float32)."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a).astype(np.complex128)
    b = np.asarray(b.detach().cpu() if isinstance(b, torch.Tensor) else b).astype(np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_kdv_gallery(dev):
    """KortewegDeVries(0.022**2, -1) on [0, 2) with N = 256, u0 = cos(pi x), dt = 1e-3, the default
    integrator, recorded every 10 steps (the reference's KdV gallery run, first 300 steps)."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.traj_recorder import AutoRecorder, IntervalController

    mi = [(0, 2.0, 256)]
    mesh = tfa.MeshGrid(mi, device=dev)
    x = mesh.bc_mesh_grid()
    u0 = torch.cos(torch.pi * x)
    kdv = pde.KortewegDeVries(0.022**2, -1)
    traj = kdv.integrate(u_0=u0, mesh=mesh, dt=0.001, step=300, trajectory_recorder=AutoRecorder(IntervalController(10)))
    ref = refpath.integrate(refpath.kdv_terms(0.022**2, -1), mi, u0.cpu(), 0.001, 300, scheme="SETDRK4",
                            record=lambda s: s % 10 == 0)
    e = _rel(traj, ref)
    print(f"KdV gallery 300 steps: rel-L2 {e:.2e}")
    assert traj.shape == ref.shape and e <= TOL


@pytest.mark.parametrize("scheme", ["ETDRK1", "ETDRK2", "SETDRK1", "SETDRK2", "SETDRK3", "SETDRK4", "RK4"])
def test_kdv_every_scheme(dev, scheme):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator, RKIntegrator, SETDRKIntegrator

    n, L = 128, 40.0
    x = torch.arange(n, dtype=torch.float64) * L / n
    u0 = torch.stack([-(c / 2) / torch.cosh(math.sqrt(c) / 2 * (x - 15.0)) ** 2 for c in (0.5, 1.0)]).reshape(2, 1, n)
    u0 = u0.float()
    mi = [(0, L, n)]
    op = pde.KortewegDeVries()
    enum = getattr(ETDRKIntegrator, scheme, None) or getattr(SETDRKIntegrator, scheme, None) or RKIntegrator[scheme]
    op.set_integrator(enum)
    dt = 2e-4 if scheme == "RK4" else 0.01
    out = op.integrate(u0.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=dt, step=30)
    ref = refpath.integrate(refpath.kdv_terms(), mi, u0, dt, 30, scheme=scheme)
    e = _rel(out, ref)
    print(f"KdV {scheme}: rel-L2 {e:.2e}")
    assert e <= TOL


def test_linear_complex_l_etdrk0_call_and_solve(dev):
    """Airy u_t = -0.3 u_xxx (1D) and advection-diffusion u_t = -0.7 u_y + 0.01 lap u (2D): ETDRK0
    with a complex E, Operator.__call__ (L u) and LinearOperator.solve (where(L == 0, 1, 1/L) u)."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.operator import Laplacian, SpatialDerivative
    from torchfsm_amd.workloads import fourier_modes_field

    cases = [(1, 128, [(-0.3, "deriv", (0, 3))], lambda: -0.3 * SpatialDerivative(0, 3)),
             (2, 64, [(-0.7, "deriv", (1, 1)), (0.01, "lap", None)],
              lambda: -0.7 * SpatialDerivative(1, 1) + 0.01 * Laplacian())]
    for dims, n, terms, make in cases:
        u0 = fourier_modes_field(n, dims, 1, 2, seed=dims)
        mi = [(0, 2 * math.pi, n)] * dims
        mesh = tfa.MeshGrid(mi, device=dev)
        op = make()
        out = op.integrate(u0.to(dev), mesh=mesh, dt=0.01, step=50)
        e1 = _rel(out, refpath.integrate(terms, mi, u0, 0.01, 50))
        # L u^ multiplies the round-off of u^ at the top modes by |L| (0.3 a^3 ~ 8e4 here): the bar
        # for one application is the reference's own float32 error against float64
        ref64 = refpath.apply(terms, mi, u0.double())
        e2 = _rel(op(u0.to(dev), mesh=mesh), ref64) / (_rel(refpath.apply(terms, mi, u0), ref64) + 1e-6)
        rm = refpath.Mesh(mi)
        Lr = refpath.build_linear(rm, terms, 1)
        sol_ref = rm.ifft(torch.where(Lr == 0, 1.0, 1 / Lr) * rm.fft(u0)).real
        e3 = _rel(op.solve(u0.to(dev)), sol_ref)
        print(f"complex L {dims}D: ETDRK0 {e1:.2e} call (error / reference's error) {e2:.2f} solve {e3:.2e}")
        assert e1 <= TOL and e2 <= 2.0 and e3 <= TOL


def test_spatial_derivative_on_a_vector_field_is_rejected(dev):
    """SpatialDerivative only takes a scalar field (_spatial_derivative.py:18-24), also beside a
    product term; the scalar case with a product term is covered in test_gpu_general.py."""
    import torchfsm_amd as tfa
    from torchfsm_amd.operator import Convection, SpatialDerivative

    mesh = tfa.MeshGrid([(0, 1, 32)] * 2, device=dev)
    op = SpatialDerivative(0, 1) + Convection()
    with pytest.raises(ValueError, match="scalar field"):
        op.integrate(torch.zeros(1, 2, 32, 32, device=dev), mesh=mesh, dt=0.01, step=1)
