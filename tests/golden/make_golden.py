"""Write golden input/output vectors from the reference torchfsm (run in the build container only).

The reference is pure Python on PyTorch; it is imported here (never shipped, never run on a
GPU) to pin the one thing no closed form, invariant or independent FFT can pin: the
reference's own float32 results -- its op order, its complex64 ETDRK closed-form coefficients
(with their float32 cancellation), its float32 contour means and the accumulated rounding of the
integrate loop.  Every fixture is small (< 1 MB) and stores its inputs next to its outputs.

Usage:  python tests/golden/make_golden.py   (writes tests/golden/*.npz)

Fixtures (reference call sites in brackets):
  g1_burgers1d      Burgers(0.01), N=128 on [0,1), u0 = sin(2 pi x)+0.5, dt 0.01, 200 steps,
                    AutoRecorder -> (1,201,1,128)            [pde.py:13-25, _base.py:676-751,
                                                              traj_recorder.py:95-125]
  g2_ns64           NavierStokesVorticity(1000, kolm_force(y)), 64^2 on [0,2pi)^2, B=2,
                    SETDRK4, dt 1e-3, 20 steps               [pde.py:66-83, field.py:128-148]
  g3_burgers3d      Burgers(0.01), 16^3 on [0,1)^3, C=3, B=1, ETDRK2 and SETDRK2, dt 5e-3, 10 steps
  g4_burgers2d_grad Burgers(nu) 32^2 on [0,1)^2, C=2, B=2, SETDRK2, dt 0.01, 10 steps; autograd
                    grads of the MSE loss w.r.t. u0 and nu (float64)
  g5_coefs          SETDRK4 coefficient arrays (NS 64^2, dt 1e-3) and complex64 ETDRK2 coefficient
                    arrays (Burgers 2D nu=0.01 on [0,1)^2 at 128^2, dt 0.01)
                                                             [_uncached.py:175-211, _etdrk.py:54-70]
  g6_nonlinear      one evaluation of VorticityConvection (64^2) and Convection (1D/2D/3D small)
                    through Operator.__call__(return_in_fourier=True)   [_base.py:753-790]
  g7_ks1d           KuramotoSivashinsky, N=512 on [0,32pi), SETDRK4, dt 0.1, 10 steps
  g8_linear         Laplacian().integrate (ETDRK0 exact linear path), 2D 32^2, 5 steps
  g9_ns128          NavierStokesVorticity(1000, kolm_force(y)), 128^2, B=1, SETDRK4, dt 1e-3,
                    100 steps (float32 and float64) -- a longer run to pin accumulated rounding
  g10_nsp_force     NavierStokes(100, force=ExplicitSource(f)), velocity form, 2D 32^2 (B=2) and
                    3D 16^3 (B=1), every ETD scheme (float64: SETDRK4), 6 steps, u0 with content at
                    every frequency;
                    and one __call__ -- pins the forced core's in-place masking of the integrator's
                    stage tensors and its double force   [_navier_stokes.py:221-280, pde.py:85-100]

Usage for one fixture:  python tests/golden/make_golden.py g10
"""

import math
import os
import sys

import numpy as np
import torch

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference():
    sys.path.insert(0, REF)
    import torchfsm  # noqa: F401
    from torchfsm.mesh import MeshGrid, FourierMesh
    from torchfsm import pde, field
    from torchfsm.operator import VorticityConvection, Convection, Laplacian, ExplicitSource
    from torchfsm.integrator import ETDRKIntegrator, SETDRKIntegrator
    from torchfsm.traj_recorder import AutoRecorder

    return dict(
        MeshGrid=MeshGrid,
        FourierMesh=FourierMesh,
        pde=pde,
        field=field,
        VorticityConvection=VorticityConvection,
        Convection=Convection,
        Laplacian=Laplacian,
        ETDRKIntegrator=ETDRKIntegrator,
        SETDRKIntegrator=SETDRKIntegrator,
        AutoRecorder=AutoRecorder,
        ExplicitSource=ExplicitSource,
    )


def ns_initial_vorticity(n, batch, seed, kmax=32.0, dtype=torch.float64):
    """Gaussian random vorticity with amplitude 1/|k| for 1<=|k|<=kmax, zero mean, unit RMS."""
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


def fourier_modes_field(n, dims, channels, batch, seed, nmodes=8, kabs=4, length=1.0):
    """Seeded sum of random Fourier modes per (batch, channel), integer k with 1<=|k|<=kabs."""
    g = torch.Generator().manual_seed(seed)
    xs = [torch.arange(n, dtype=torch.float64) * length / n for _ in range(dims)]
    grids = torch.meshgrid(*xs, indexing="ij")
    out = torch.zeros(batch, channels, *([n] * dims), dtype=torch.float64)
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
    return out


def g1(R):
    out = {}
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        mesh = R["MeshGrid"]([(0, 1, 128)], dtype=dt_)
        x = mesh.bc_mesh_grid()
        u0 = torch.sin(2 * torch.pi * x) + 0.5
        op = R["pde"].Burgers(0.01)
        traj = op.integrate(u0, mesh=mesh, dt=0.01, step=200, trajectory_recorder=R["AutoRecorder"]())
        out[f"u0_{name}"] = u0.numpy()
        out[f"traj_{name}"] = traj.numpy()
    np.savez_compressed(os.path.join(HERE, "g1_burgers1d.npz"), **out)


def g2(R):
    out = {}
    w0 = ns_initial_vorticity(64, 2, seed=1234)
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        mesh = R["MeshGrid"]([(0, 2 * math.pi, 64)] * 2, dtype=dt_)
        y = mesh.bc_mesh_grid()[1]
        op = R["pde"].NavierStokesVorticity(1000, R["field"].kolm_force(y))
        op.set_integrator(R["SETDRKIntegrator"].SETDRK4)
        u0 = w0.to(dt_)
        res = op.integrate(u0, mesh=mesh, dt=1e-3, step=20)
        out[f"u0_{name}"] = u0.numpy()
        out[f"out_{name}"] = res.numpy()
    np.savez_compressed(os.path.join(HERE, "g2_ns64.npz"), **out)


def g3(R):
    out = {}
    u0 = fourier_modes_field(16, 3, 3, 1, seed=7)
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        mesh = R["MeshGrid"]([(0, 1, 16)] * 3, dtype=dt_)
        for sname, scheme in (
            ("etdrk2", R["ETDRKIntegrator"].ETDRK2),
            ("setdrk2", R["SETDRKIntegrator"].SETDRK2),
        ):
            op = R["pde"].Burgers(0.01)
            op.set_integrator(scheme)
            res = op.integrate(u0.to(dt_), mesh=mesh, dt=5e-3, step=10)
            out[f"out_{sname}_{name}"] = res.numpy()
        out[f"u0_{name}"] = u0.to(dt_).numpy()
    np.savez_compressed(os.path.join(HERE, "g3_burgers3d.npz"), **out)


def g4(R):
    out = {}
    u0 = fourier_modes_field(32, 2, 2, 2, seed=11)
    target = fourier_modes_field(32, 2, 2, 2, seed=12)
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        mesh = R["MeshGrid"]([(0, 1, 32)] * 2, dtype=dt_)
        nu = torch.tensor(0.01, dtype=dt_, requires_grad=True)
        u = u0.to(dt_).clone().requires_grad_(True)
        op = R["pde"].Burgers(nu)
        op.set_integrator(R["SETDRKIntegrator"].SETDRK2)
        res = op.integrate(u, mesh=mesh, dt=0.01, step=10)
        loss = ((res - target.to(dt_)) ** 2).mean()
        loss.backward()
        out[f"u0_{name}"] = u0.to(dt_).numpy()
        out[f"target_{name}"] = target.to(dt_).numpy()
        out[f"out_{name}"] = res.detach().numpy()
        out[f"loss_{name}"] = np.array(loss.item())
        out[f"grad_u0_{name}"] = u.grad.numpy()
        out[f"grad_nu_{name}"] = np.array(nu.grad.item())
    np.savez_compressed(os.path.join(HERE, "g4_burgers2d_grad.npz"), **out)


def g5(R):
    out = {}
    # SETDRK4 coefficients for NS 64^2 (L = lap/Re - 0.1), dt 1e-3
    mesh = R["MeshGrid"]([(0, 2 * math.pi, 64)] * 2)
    y = mesh.bc_mesh_grid()[1]
    op = R["pde"].NavierStokesVorticity(1000, R["field"].kolm_force(y))
    op.set_integrator(R["SETDRKIntegrator"].SETDRK4)
    u0 = ns_initial_vorticity(64, 1, seed=3).float()
    op.integrate(u0, mesh=mesh, dt=1e-3, step=1)
    it = op._state_dict["integrator"]
    out["ns_L"] = op._state_dict["linear_coef"].numpy()
    out["ns_E"] = it._exp_term.numpy()
    out["ns_E2"] = it._half_exp_term.numpy()
    for i in (1, 4, 5, 6):
        out[f"ns_c{i}"] = getattr(it, f"_coef_{i}").numpy()
    # complex64 ETDRK2 coefficients, Burgers 2D nu=0.01 on [0,1)^2, 128^2, dt 0.01
    mesh = R["MeshGrid"]([(0, 1, 128)] * 2)
    op = R["pde"].Burgers(0.01)
    op.set_integrator(R["ETDRKIntegrator"].ETDRK2)
    u0 = fourier_modes_field(128, 2, 2, 1, seed=5).float()
    op.integrate(u0, mesh=mesh, dt=0.01, step=1)
    it = op._state_dict["integrator"]
    out["b2_L"] = op._state_dict["linear_coef"].numpy()
    out["b2_E"] = it._exp_term.numpy()
    out["b2_c1"] = it._coef_1.numpy()
    out["b2_c2"] = it._coef_2.numpy()
    np.savez_compressed(os.path.join(HERE, "g5_coefs.npz"), **out)


def g6(R):
    out = {}
    FM = R["FourierMesh"]
    # VorticityConvection at 64^2 on [0, 2pi)^2
    w = ns_initial_vorticity(64, 2, seed=21).float()
    mesh = FM([(0, 2 * math.pi, 64)] * 2)
    vc = R["VorticityConvection"]()
    out["vc_in"] = w.numpy()
    out["vc_out"] = vc(u=w, mesh=mesh, return_in_fourier=True).numpy()
    # Convection 1D (N=64 on [0,1)), 2D (32^2 on [0,1)^2, C=2), 3D (16^3 on [0,1)^3, C=3)
    for tag, n, d, c in (("c1", 64, 1, 1), ("c2", 32, 2, 2), ("c3", 16, 3, 3)):
        u = fourier_modes_field(n, d, c, 2, seed=30 + d).float()
        mesh = FM([(0, 1, n)] * d)
        conv = R["Convection"]()
        out[f"{tag}_in"] = u.numpy()
        out[f"{tag}_out"] = conv(u=u, mesh=mesh, return_in_fourier=True).numpy()
    np.savez_compressed(os.path.join(HERE, "g6_nonlinear.npz"), **out)


def g7(R):
    out = {}
    length = 32 * math.pi
    mesh = R["MeshGrid"]([(0, length, 512)])
    x = mesh.bc_mesh_grid()
    u0 = torch.cos(x / 16) * (1 + torch.sin(x / 16))
    op = R["pde"].KuramotoSivashinsky()
    op.set_integrator(R["SETDRKIntegrator"].SETDRK4)
    res = op.integrate(u0, mesh=mesh, dt=0.1, step=10)
    out["u0"] = u0.numpy()
    out["out"] = res.numpy()
    np.savez_compressed(os.path.join(HERE, "g7_ks1d.npz"), **out)


def g8(R):
    out = {}
    u0 = fourier_modes_field(32, 2, 1, 2, seed=41).float()
    mesh = R["MeshGrid"]([(0, 1, 32)] * 2)
    op = 0.01 * R["Laplacian"]()
    res = op.integrate(u0, mesh=mesh, dt=0.05, step=5)
    out["u0"] = u0.numpy()
    out["out"] = res.numpy()
    np.savez_compressed(os.path.join(HERE, "g8_linear.npz"), **out)


def g9(R):
    out = {}
    w0 = ns_initial_vorticity(128, 1, seed=99)
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        mesh = R["MeshGrid"]([(0, 2 * math.pi, 128)] * 2, dtype=dt_)
        y = mesh.bc_mesh_grid()[1]
        op = R["pde"].NavierStokesVorticity(1000, R["field"].kolm_force(y))
        u0 = w0.to(dt_)
        res = op.integrate(u0, mesh=mesh, dt=1e-3, step=100)
        out[f"u0_{name}"] = u0.numpy()
        out[f"out_{name}"] = res.numpy()
    np.savez_compressed(os.path.join(HERE, "g9_ns128.npz"), **out)


def g10(R):
    out = {}
    gen = torch.Generator().manual_seed(10)
    for dims, n, batch in ((2, 32, 2), (3, 16, 1)):
        u0 = 0.5 * fourier_modes_field(n, dims, dims, batch, seed=100 + dims)
        u0 = u0 + 0.05 * torch.randn(u0.shape, generator=gen, dtype=torch.float64)
        xs = torch.meshgrid(*[torch.arange(n, dtype=torch.float64) / n] * dims, indexing="ij")
        f = torch.zeros(1, dims, *([n] * dims), dtype=torch.float64)
        f[0, 0] = 0.5 * torch.sin(2 * math.pi * (xs[0] + 2 * xs[1]))
        f[0, 1] = 0.3 * torch.cos(2 * math.pi * 3 * xs[0])
        if dims == 3:
            f[0, 2] = 0.2 * torch.sin(2 * math.pi * (xs[2] - xs[1]))
        tag = f"{dims}d"
        out[f"u0_{tag}"] = u0.numpy()
        out[f"force_{tag}"] = f.numpy()
        for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
            mesh = R["MeshGrid"]([(0, 1, n)] * dims, dtype=dt_)
            schemes = [("etdrk1", R["ETDRKIntegrator"].ETDRK1), ("etdrk2", R["ETDRKIntegrator"].ETDRK2),
                       ("setdrk1", R["SETDRKIntegrator"].SETDRK1), ("setdrk2", R["SETDRKIntegrator"].SETDRK2),
                       ("setdrk3", R["SETDRKIntegrator"].SETDRK3), ("setdrk4", R["SETDRKIntegrator"].SETDRK4)]
            for sname, scheme in schemes:
                if name == "f64" and sname != "setdrk4":  # float64 only for the default scheme (size)
                    continue
                op = R["pde"].NavierStokes(100, force=R["ExplicitSource"](f.to(dt_)))
                op.set_integrator(scheme)
                res = op.integrate(u0.to(dt_), mesh=mesh, dt=2e-3, step=6)
                out[f"out_{tag}_{sname}_{name}"] = res.numpy()
            op = R["pde"].NavierStokes(100, force=R["ExplicitSource"](f.to(dt_)))
            out[f"call_{tag}_{name}"] = op(u0.to(dt_), mesh=mesh, return_in_fourier=True).numpy()
    np.savez_compressed(os.path.join(HERE, "g10_nsp_force.npz"), **out)


def main():
    torch.set_num_threads(min(8, os.cpu_count() or 1))
    R = _import_reference()
    fns = (g1, g2, g3, g4, g5, g6, g7, g8, g9, g10)
    only = set(sys.argv[1:])
    for fn in fns:
        if only and fn.__name__ not in only:
            continue
        fn(R)
        print("wrote", fn.__name__, flush=True)


if __name__ == "__main__":
    main()
