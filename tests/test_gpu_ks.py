"""KSConvection (1/2 |grad phi|^2, _ks_convection.py:8-67) and pde.KuramotoSivashinskyHighDim
(pde.py:41-49) on the engine against the oracle on the same inputs, in 1D, 2D and 3D, with and
without remove_mean.  With remove_mean the reference subtracts the mean over the whole batch, so
the batches here hold several samples of different amplitude.  Tolerance: rel-L2 <= 1e-5 (float32)."""

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


def _phi0(n, dims, batch, seed):
    from torchfsm_amd.workloads import fourier_modes_field

    u = fourier_modes_field(n, dims, 1, batch, seed=seed)
    amp = torch.linspace(0.5, 1.5, batch).reshape((batch,) + (1,) * (dims + 1))
    return (u * amp).float()


@pytest.mark.parametrize("dims,n,batch,steps,L,dt", [(1, 128, 4, 40, 22.0, 0.05), (1, 100, 3, 40, 60.0, 0.1),
                                                     (2, 64, 3, 15, 22.0, 0.05), (2, 100, 2, 10, 60.0, 0.1),
                                                     (3, 32, 2, 5, 22.0, 0.05), (3, 48, 1, 3, 30.0, 0.05)])
@pytest.mark.parametrize("remove_mean", [True, False])
def test_ks_highdim(dev, dims, n, batch, steps, L, dt, remove_mean):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.operator import Biharmonic, KSConvection, Laplacian
    from torchfsm_amd.traj_recorder import AutoRecorder, IntervalController

    u0 = _phi0(n, dims, batch, seed=10 * dims + n)
    mi = [(0, L, n)] * dims
    op = -Laplacian() - Biharmonic() - KSConvection(remove_mean)
    rec = AutoRecorder(IntervalController(max(1, steps // 4)))
    traj = op.integrate(u0.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=dt, step=steps, trajectory_recorder=rec)
    pred = rec.control_func
    ref = refpath.integrate(refpath.ks_highdim_terms(remove_mean), mi, u0, dt, steps, scheme="SETDRK4",
                            record=lambda s: bool(pred(s)))
    e = _rel(traj, ref)
    # the per-sample means drift apart only through the batch mean: check the DC of the last frame
    dc = traj[:, -1].mean(dim=tuple(range(1, dims + 2)))
    dcr = ref[:, -1].mean(dim=tuple(range(1, dims + 2)))
    print(f"KS {dims}D {n} B={batch} remove_mean={remove_mean}: rel-L2 {e:.2e}; DC {dc.cpu().numpy()} "
          f"ref {dcr.numpy()}")
    assert traj.shape == ref.shape and e <= TOL
    assert torch.allclose(dc.cpu(), dcr, atol=1e-5 * float(ref.abs().max()))


@pytest.mark.parametrize("dims,n", [(1, 128), (2, 64), (3, 32)])
def test_ks_convection_call_and_schemes(dev, dims, n):
    """Operator.__call__ of KuramotoSivashinskyHighDim (one evaluation, L u + N(u)) and the ETDRK2
    and RK4 integrators on the same term.  The domain (L = 200) keeps the largest |L| = k^4 - k^2
    of the mesh moderate: L u^ multiplies the float32 round-off of u^ at the top modes by |L|, and
    on the L = 22 domain (k^4 ~ 1e5) that amplified noise alone exceeds the tolerance."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator, RKIntegrator

    u0 = _phi0(n, dims, 2, seed=7 + dims)
    mi = [(0, 200.0, n)] * dims
    mesh = tfa.MeshGrid(mi, device=dev)
    op = pde.KuramotoSivashinskyHighDim()
    out = op(u0.to(dev), mesh=mesh)
    ref = refpath.apply(refpath.ks_highdim_terms(), mi, u0)
    # L u^ carries |L| times the float32 round-off of u^ at the top modes, so the bar for one
    # evaluation is the reference's own float32 error against float64 (plus 1e-6)
    ref64 = refpath.apply(refpath.ks_highdim_terms(), mi, u0.double())
    e0 = _rel(out, ref64) / (_rel(ref, ref64) + 1e-6)
    op.set_integrator(ETDRKIntegrator.ETDRK2)
    e1 = _rel(op.integrate(u0.to(dev), mesh=mesh, dt=0.05, step=10),
              refpath.integrate(refpath.ks_highdim_terms(), mi, u0, 0.05, 10, scheme="ETDRK2"))
    op.set_integrator(RKIntegrator.RK4)
    e2 = _rel(op.integrate(u0.to(dev), mesh=mesh, dt=0.05, step=10),
              refpath.integrate(refpath.ks_highdim_terms(), mi, u0, 0.05, 10, scheme="RK4"))
    print(f"KS {dims}D call (error / reference's error) {e0:.2f}  ETDRK2 {e1:.2e}  RK4 {e2:.2e}")
    assert e0 <= 2.0 and e1 <= TOL and e2 <= TOL


def test_ks_rejects_sub_batches_and_vectors(dev):
    import torchfsm_amd as tfa
    from torchfsm_amd import pde
    from torchfsm_amd.operator import KSConvection

    mesh = tfa.MeshGrid([(0, 22.0, 64)] * 2, device=dev)
    u0 = _phi0(64, 2, 16, seed=3).to(dev)
    op = pde.KuramotoSivashinskyHighDim()
    op.integrate(u0, mesh=mesh, dt=0.01, step=1)
    plan = op._state_dict["integrator"]
    assert plan.tuning[4] == 0
    with pytest.raises(ValueError, match="mean"):
        plan.set_tuning(col_tile=16, sub_batch=8)
    with pytest.raises(ValueError, match="mean"):
        plan.set_streams(2)
    with pytest.raises(NotImplementedError, match="scalar"):
        KSConvection().integrate(torch.zeros(1, 2, 64, 64, device=dev), mesh=mesh, dt=0.01, step=1)
