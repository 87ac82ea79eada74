"""Public API paths of OperatorLike on the engine against the oracle on the same inputs:
trajectory recorders (interval, start, include_initial_state, CPU), return_in_fourier, u_0_fft
input, every ETD scheme and every explicit Runge-Kutta tableau in 1D/2D/3D, per-sample coefficients, LinearOperator.solve,
Operator.__call__ in 3D, progressive integration, HIP-graph replay and per-sample domain lengths.
Tolerance: rel-L2 <= 1e-5 in float32 (_base.py:217-262, 676-790; traj_recorder.py:7-148)."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a)
    b = np.asarray(b.detach().cpu() if isinstance(b, torch.Tensor) else b)
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a, b = a.astype(np.complex128), b.astype(np.complex128)
    else:
        a, b = a.astype(np.float64), b.astype(np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _ns(dev, n=64):
    import torchfsm_amd as tfa
    from torchfsm_amd import field, pde

    mesh = tfa.MeshGrid([(0, 2 * math.pi, n)] * 2, device=dev)
    y = mesh.bc_mesh_grid()[1]
    return pde.NavierStokesVorticity(1000, field.kolm_force(y)), mesh, y


def test_recorders_interval_start_initial_state(dev):
    from oracle import refpath
    from torchfsm_amd.traj_recorder import AutoRecorder, CPURecorder, IntervalController
    from torchfsm_amd.workloads import ns_initial_vorticity

    op, mesh, y = _ns(dev)
    w0 = ns_initial_vorticity(64, 2, seed=5)
    terms = refpath.ns_terms(1000, y.cpu())
    mi = [(0, 2 * math.pi, 64)] * 2
    for ctrl, init in ((IntervalController(3, 2), True), (IntervalController(2), False), (None, True)):
        rec = AutoRecorder(ctrl, include_initial_state=init)
        traj = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=11, trajectory_recorder=rec)
        pred = rec.control_func
        ref = refpath.integrate(terms, mi, w0, 1e-3, 11, scheme="SETDRK4", record=lambda s: bool(pred(s)))
        assert tuple(traj.shape) == tuple(ref.shape), (traj.shape, ref.shape)
        e = _rel(traj, ref)
        print(f"recorder ctrl={ctrl and (ctrl.interval, ctrl.start)} init={init}: T={traj.shape[1]} rel-L2 {e:.2e}")
        assert e <= TOL
    crec = CPURecorder(IntervalController(4))
    ctraj = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=8, trajectory_recorder=crec)
    assert ctraj.device.type == "cpu"
    arec = AutoRecorder(IntervalController(4))
    atraj = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=8, trajectory_recorder=arec)
    assert torch.equal(ctraj, atraj.cpu())
    # trajectory in Fourier space: full complex spectra of the recorded frames
    frec = AutoRecorder(IntervalController(4))
    ftraj = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=8, trajectory_recorder=frec, return_in_fourier=True)
    fref = refpath.integrate(terms, mi, w0, 1e-3, 8, scheme="SETDRK4", record=lambda s: s % 4 == 0,
                             return_in_fourier=True)
    assert ftraj.dtype == torch.complex64 and _rel(ftraj, fref) <= TOL


def test_return_in_fourier_and_u0_fft(dev):
    from oracle import refpath
    from torchfsm_amd.workloads import ns_initial_vorticity

    op, mesh, y = _ns(dev)
    w0 = ns_initial_vorticity(64, 2, seed=6)
    terms = refpath.ns_terms(1000, y.cpu())
    mi = [(0, 2 * math.pi, 64)] * 2
    out = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=10, return_in_fourier=True)
    ref = refpath.integrate(terms, mi, w0, 1e-3, 10, scheme="SETDRK4", return_in_fourier=True)
    assert out.dtype == torch.complex64 and tuple(out.shape) == (2, 1, 64, 64)
    e1 = _rel(out, ref)
    w0f = torch.fft.fftn(w0, dim=(-2, -1))
    out2 = op.integrate(u_0_fft=w0f.to(dev), mesh=mesh, dt=1e-3, step=10)
    ref2 = refpath.integrate(terms, mi, None, 1e-3, 10, scheme="SETDRK4", u0_fft=w0f)
    e2 = _rel(out2, ref2)
    print(f"return_in_fourier {e1:.2e}; u_0_fft input {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL


RK_NAMES = ["Euler", "Midpoint", "Heun12", "Ralston12", "BogackiShampine23", "RK4", "RK4_38Rule", "Dorpi45",
            "Fehlberg45", "CashKarp45"]


@pytest.mark.parametrize("name", RK_NAMES)
def test_every_rk_integrator_1d_2d_3d(dev, name):
    """Explicit Runge-Kutta stepping (_rk.py:43-58) on the engine's L u + N(u), every tableau."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import RKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    for dims, n, steps in ((1, 256, 20), (2, 64, 10), (3, 32, 5)):
        u = 0.5 * fourier_modes_field(n, dims, dims, 2, seed=50 + dims)
        mi = [(0, 1, n)] * dims
        op = pde.Burgers(0.01)
        op.set_integrator(RKIntegrator[name])
        dt = 2e-4 if dims == 1 else 5e-4
        out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=dt, step=steps)
        ref = refpath.integrate(refpath.burgers_terms(0.01), mi, u, dt, steps, scheme=name)
        e = _rel(out, ref)
        print(f"{name} {dims}D {n}: rel-L2 {e:.2e}")
        assert e <= TOL


def test_rk_forced_ns_recorder_and_options(dev):
    """RK4 on forced 2D Navier-Stokes (vorticity form, drag and Kolmogorov forcing) with an
    AutoRecorder, return_in_fourier, and the option checks of set_integrator."""
    from oracle import refpath
    from torchfsm_amd.integrator import RKIntegrator
    from torchfsm_amd.traj_recorder import AutoRecorder, IntervalController
    from torchfsm_amd.workloads import ns_initial_vorticity

    op, mesh, y = _ns(dev)
    op.set_integrator(RKIntegrator.RK4)
    w0 = ns_initial_vorticity(64, 2, seed=9)
    terms = refpath.ns_terms(1000, y.cpu())
    mi = [(0, 2 * math.pi, 64)] * 2
    rec = AutoRecorder(IntervalController(3))
    traj = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=10, trajectory_recorder=rec)
    ref = refpath.integrate(terms, mi, w0, 1e-3, 10, scheme="RK4", record=lambda s: s % 3 == 0)
    assert tuple(traj.shape) == tuple(ref.shape)
    e1 = _rel(traj, ref)
    out = op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=10, return_in_fourier=True)
    e2 = _rel(out, refpath.integrate(terms, mi, w0, 1e-3, 10, scheme="RK4", return_in_fourier=True))
    print(f"RK4 NS recorder {e1:.2e}; fourier {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL
    op.set_integrator(RKIntegrator.Dorpi45, adaptive=True)
    with pytest.raises(ValueError, match="adaptive"):
        op.integrate(w0.to(dev), mesh=mesh, dt=1e-3, step=2)


@pytest.mark.parametrize("scheme", ["ETDRK1", "ETDRK2", "SETDRK1", "SETDRK2", "SETDRK3", "SETDRK4"])
def test_every_scheme_1d_2d_3d(dev, scheme):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator, SETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    enum = getattr(ETDRKIntegrator, scheme, None) or getattr(SETDRKIntegrator, scheme)
    for dims, n, steps in ((1, 256, 20), (2, 64, 10), (3, 32, 5)):
        u = 0.5 * fourier_modes_field(n, dims, dims, 2, seed=40 + dims)
        mi = [(0, 1, n)] * dims
        op = pde.Burgers(0.01)
        op.set_integrator(enum)
        out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=2e-3, step=steps)
        ref = refpath.integrate(refpath.burgers_terms(0.01), mi, u, 2e-3, steps, scheme=scheme)
        e = _rel(out, ref)
        print(f"{scheme} {dims}D {n}: rel-L2 {e:.2e}")
        assert e <= TOL


@pytest.mark.parametrize("dims,n", [(2, 64), (3, 32)])
def test_per_sample_viscosity(dev, dims, n):
    """Burgers with a [B,1,...] viscosity tensor (per-sample coefficient tables, SETDRK4): each
    sample matches its own scalar-viscosity run."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import SETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    nus = [0.005, 0.01, 0.03]
    u = 0.5 * fourier_modes_field(n, dims, dims, 3, seed=50)
    op = pde.Burgers(torch.tensor(nus).reshape((3,) + (1,) * (dims + 1)))
    op.set_integrator(SETDRKIntegrator.SETDRK4)
    mi = [(0, 1, n)] * dims
    out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=2e-3, step=15)
    for b, nu in enumerate(nus):
        ref = refpath.integrate(refpath.burgers_terms(nu), mi, u[b:b + 1], 2e-3, 15, scheme="SETDRK4")
        e = _rel(out[b:b + 1], ref)
        print(f"sample {b} nu={nu}: rel-L2 {e:.2e}")
        assert e <= TOL


def test_solve_and_call_3d(dev):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.workloads import fourier_modes_field

    mi = [(0, 1, 32)] * 3
    u = fourier_modes_field(32, 3, 3, 2, seed=60)
    # Operator.__call__: L u^ + N(u^) of 3D Burgers
    op = pde.Burgers(0.02)
    got = op(u=u.to(dev), mesh=tfa.FourierMesh(mi, device=dev), return_in_fourier=True)
    ref = refpath.apply(refpath.burgers_terms(0.02), mi, u, return_in_fourier=True)
    e1 = _rel(got, ref)
    # LinearOperator.solve: b^ / L with 1/0 := 1
    lin = 0.5 * tfa.Laplacian() - 0.1 * tfa.Biharmonic()   # a LinearOperator (L = 0 at k = 0)
    assert isinstance(lin, tfa.LinearOperator)
    got2 = lin.solve(b=u.to(dev), mesh=tfa.FourierMesh(mi, device=dev), n_channel=3)
    m = refpath.Mesh(mi)
    L = refpath.build_linear(m, [(0.5, "lap", None), (-0.1, "biharm", None)], 3)
    ref2 = m.ifft(m.fft(u) * torch.where(L == 0, 1.0, 1 / L)).real
    e2 = _rel(got2, ref2)
    print(f"3D __call__ {e1:.2e}; solve {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL


def test_progressive_and_graph_replay(dev):
    from torchfsm_amd.workloads import ns_initial_vorticity

    op, mesh, y = _ns(dev)
    w0 = ns_initial_vorticity(64, 2, seed=7).to(dev)
    a = op.integrate(w0, mesh=mesh, dt=1e-3, step=30)
    b = op.integrate(w0, mesh=mesh, dt=1e-3, step=30, progressive=True)
    assert torch.equal(a, b)
    # a captured graph of 5 steps, replayed 6 times with eager work and syncs in between,
    # equals 30 eager steps after every replay
    plan = op._prepare(2, dev, 1e-3)
    ug = plan.forward(w0)
    ue = ug.clone()
    g = plan.graph(ug, 5)
    for r in range(6):
        g.launch()
        plan.step(ue, 5)
        torch.cuda.synchronize()
        _ = (ue.abs().sum() + 1).item()  # eager work between replays
        assert torch.equal(ug, ue), f"replay {r}"
    g.close()


def test_ks_per_sample_lengths(dev):
    """1D KS plan with a domain length per sample (serving of mixed requests) against the oracle
    run per length."""
    from oracle import refpath
    from torchfsm_amd.engine import SpectralPlan

    n, lens = 512, [32 * math.pi, 60 * math.pi, 100 * math.pi]
    x = torch.arange(n, dtype=torch.float64) / n
    u = torch.stack([torch.cos(2 * math.pi * 4 * x) + 0.3 * torch.sin(2 * math.pi * 7 * x + b) for b in range(3)])
    u = u.reshape(3, 1, n).float()
    plan = SpectralPlan([(0, lens[0], n)], 3, 1, nonlinear="conv", nonlinear_coef=-1.0,
                        linear=[("lap", [-1.0] * 3), ("biharm", [-1.0] * 3)], device=dev, lengths_per_sample=lens)
    plan.set_scheme("SETDRK4", 0.1)
    uh = plan.forward(u.to(dev))
    plan.step(uh, 20)
    out = plan.inverse(uh)
    for b, L in enumerate(lens):
        ref = refpath.integrate(refpath.ks_terms(), [(0, L, n)], u[b:b + 1], 0.1, 20, scheme="SETDRK4")
        e = _rel(out[b:b + 1], ref)
        print(f"KS L={L / math.pi:.0f}pi: rel-L2 {e:.2e}")
        assert e <= TOL
    # re-targeting a slot to another length gives that length's trajectory
    plan.set_sample(1, [-1.0, -1.0], 45 * math.pi)
    uh = plan.forward(u.to(dev))
    plan.step(uh, 20)
    ref = refpath.integrate(refpath.ks_terms(), [(0, 45 * math.pi, n)], u[1:2], 0.1, 20, scheme="SETDRK4")
    assert _rel(plan.inverse(uh)[1:2], ref) <= TOL


def test_linear_etdrk0_3d(dev):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.workloads import fourier_modes_field

    mi = [(0, 1, 32), (0, 2, 16), (0, 1, 64)]
    u = fourier_modes_field((32, 16, 64), 3, 1, 2, seed=70)
    op = 0.01 * tfa.Laplacian()
    out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=0.05, step=7)
    ref = refpath.integrate([(0.01, "lap", None)], mi, u, 0.05, 7)
    e = _rel(out, ref)
    print(f"ETDRK0 3D: rel-L2 {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("n,dims", [(64, 2), (128, 2), (32, 3), (64, 3)])
def test_navier_stokes_velocity_form(dev, n, dims):
    """pde.NavierStokes(Re) = NSPressureConvection() + Laplacian()/Re (pde.py:85-100) in 2D and 3D,
    every launch shape of the plan, against the oracle; and one direct evaluation of the term."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import autotune, pde
    from torchfsm_amd.workloads import fourier_modes_field

    mi = [(0, 2 * math.pi, n)] * dims
    u = fourier_modes_field(n, dims, dims, 2, seed=80 + n, length=2 * math.pi)
    op = pde.NavierStokes(100.0)
    out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=2e-3, step=10)
    ref = refpath.integrate(refpath.ns_velocity_terms(100.0), mi, u, 2e-3, 10, scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"NavierStokes {dims}D {n}: rel-L2 {e:.2e}")
    assert e <= TOL
    plan = op._prepare(2, dev, 2e-3)
    for cand in autotune.candidates(plan):
        plan.set_tuning(*cand)
        uh = plan.forward(u.to(dev))
        plan.step(uh, 10)
        e1 = _rel(plan.inverse(uh), ref)
        assert e1 <= TOL, (cand, e1)
    core = tfa.NSPressureConvection()
    got = core(u=u.to(dev), mesh=tfa.FourierMesh(mi, device=dev), return_in_fourier=True)
    want = refpath.apply([(1, "nsp", None)], mi, u, return_in_fourier=True)
    e2 = _rel(got, want)
    print(f"  direct NSPressureConvection rel-L2 {e2:.2e}")
    assert e2 <= TOL
    with pytest.raises(ValueError, match="external force"):
        pde.NavierStokes(100.0, force=tfa.ImplicitSource()).register_mesh(tfa.FourierMesh(mi, device=dev), dims)


def test_cpu_recorder_streams_frames_while_stepping(dev):
    """CPURecorder: frames go to pinned host memory in chunks on a copy stream while the next chunk
    steps; the result equals AutoRecorder's bit for bit (NS 2D, several ring chunks, physical and
    Fourier output, a last frame before the final step, and a schedule that records nothing)."""
    import time

    from torchfsm_amd.traj_recorder import AutoRecorder, CPURecorder, IntervalController
    from torchfsm_amd.workloads import ns_initial_vorticity

    op, mesh, y = _ns(dev)
    w0 = ns_initial_vorticity(64, 3, seed=11).to(dev)
    frame_bytes = 3 * 64 * 48 * 8  # B * spectrum (pitch 48) * complex64
    for ring_frames, ctrl, fourier in ((2, IntervalController(3), False), (5, IntervalController(2, 1), True),
                                       (1000, IntervalController(1), False)):
        op._stream_ring_bytes = ring_frames * frame_bytes
        crec = CPURecorder(ctrl)
        ct = op.integrate(w0, mesh=mesh, dt=1e-3, step=20, trajectory_recorder=crec, return_in_fourier=fourier)
        arec = AutoRecorder(ctrl)
        at = op.integrate(w0, mesh=mesh, dt=1e-3, step=20, trajectory_recorder=arec, return_in_fourier=fourier)
        print(f"ring {ring_frames} frames, fourier={fourier}: T={ct.shape[1]}, pinned={ct.is_pinned()}")
        assert ct.device.type == "cpu" and ct.is_pinned() and torch.equal(ct, at.cpu())
    op._stream_ring_bytes = 2 * frame_bytes
    none = op.integrate(w0, mesh=mesh, dt=1e-3, step=5, trajectory_recorder=CPURecorder(lambda s: False))
    assert none is None
    # overlap: 400 steps of 256^2 x 16 recorded every step (1.7 GB of frames) with the default ring
    from torchfsm_amd import field, pde
    import torchfsm_amd as tfa

    mesh2 = tfa.MeshGrid([(0, 2 * math.pi, 256)] * 2, device=dev)
    op2 = pde.NavierStokesVorticity(1000, field.kolm_force(mesh2.bc_mesh_grid()[1]))
    w2 = ns_initial_vorticity(256, 16, seed=12).to(dev)
    op2.integrate(w2, mesh=mesh2, dt=1e-3, step=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op2.integrate(w2, mesh=mesh2, dt=1e-3, step=400)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    tr = op2.integrate(w2, mesh=mesh2, dt=1e-3, step=400, trajectory_recorder=CPURecorder())
    t_stream = time.perf_counter() - t0
    # the same frames kept on the device, transformed, then copied to pinned host memory in one go
    t0 = time.perf_counter()
    arec = AutoRecorder()
    op2.integrate(w2, mesh=mesh2, dt=1e-3, step=400, trajectory_recorder=arec)
    dev_traj = arec.trajectory
    host = torch.empty(dev_traj.shape, dtype=dev_traj.dtype, pin_memory=True)
    host.copy_(dev_traj)
    t_serial = time.perf_counter() - t0
    print(f"400 steps 256^2 B=16: {t_plain * 1e3:.1f} ms without frames; {tr.numel() * 4 / 1e9:.2f} GB of frames to "
          f"host: {t_stream * 1e3:.1f} ms streamed during stepping, {t_serial * 1e3:.1f} ms recorded then copied")
    assert tr.shape == (16, 401, 1, 256, 256) and torch.equal(tr, host)


def test_plan_rejects_mismatched_tensors(dev):
    """Every SpectralPlan call checks device, dtype, contiguity and shape before a pointer reaches the
    C ABI, and the C ABI rejects a batch count below one on its own; nothing is launched."""
    import ctypes

    from torchfsm_amd._lib import lib
    from torchfsm_amd.engine import SpectralPlan

    plan = SpectralPlan([(0, 1, 32)] * 2, 2, 2, nonlinear="conv", linear=[("lap", 0.01)], device=dev)
    plan.set_scheme("SETDRK2", 1e-3)
    u = torch.zeros(2, 2, 32, 32, device=dev)
    uh = plan.forward(u)
    with pytest.raises(ValueError, match="shape"):
        plan.forward(torch.zeros(2, 2, 32, 16, device=dev))
    with pytest.raises(RuntimeError, match="GPU only"):
        plan.forward(torch.zeros(2, 2, 32, 32))
    with pytest.raises(ValueError, match="samples"):
        plan.step(uh[:1].contiguous(), 1)
    with pytest.raises(TypeError, match="complex64"):
        plan.apply(uh.real.contiguous())
    with pytest.raises(ValueError, match="contiguous"):
        plan.solve(uh.transpose(2, 3))
    with pytest.raises(ValueError, match="nbatch"):
        plan.inverse(uh, nbatch=3)
    with pytest.raises(ValueError, match="record slots"):
        plan.step(uh, 4, torch.zeros(2, dtype=torch.int32, device=dev),
                  torch.zeros((2, 1) + plan.spec_shape[1:], dtype=torch.complex64, device=dev))
    # slot values outside -1..T-1 would index past the trajectory buffer: rejected on the host,
    # from the device schedule and from the caller's host copy alike
    traj2 = torch.zeros((2, 2) + plan.spec_shape[1:], dtype=torch.complex64, device=dev)
    for bad in ([0, 1, 2], [0, -2, 1]):
        with pytest.raises(ValueError, match="record slots must lie in -1..1"):
            plan.step(uh, 2, torch.tensor(bad, dtype=torch.int32, device=dev), traj2)
        with pytest.raises(ValueError, match="record slots must lie in -1..1"):
            plan.step(uh, 2, torch.tensor([0, -1, 1], dtype=torch.int32, device=dev), traj2, rec_slot_host=bad)
    with pytest.raises(ValueError, match="channels"):
        plan.spectral_map(0, uh, 3)
    s = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    for nb in (0, -1):
        rc = lib.tfsm_forward(plan.handle, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(uh.data_ptr()), nb, s)
        assert rc != 0 and b"nbatch" in lib.tfsm_last_error()
    # the plan still works after the rejections
    plan.step(uh, 2)
    assert torch.isfinite(plan.inverse(uh)).all()
