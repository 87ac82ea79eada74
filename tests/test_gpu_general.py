"""Any mesh size and the term combinations without fused kernels (k_gfft.hip, k_general.hip)
against numpy.fft and the oracle: smooth lengths, prime lengths (Bluestein), lengths beyond 4096
points (four-step), 2D 1024^2 and 3D 256^3 meshes, ConservativeConvection in 3D, a complex linear
symbol with a product term in 2D / 3D, and the unfused path forced on meshes the fused kernels
cover.  The reference transforms and steps any mesh (mesh.py:481-491, _base.py:676-751).
Tolerances: transforms rel-L2 <= 2e-6 (direct) / 1e-5 (Bluestein, four-step) against float64
numpy; stepping rel-L2 <= 1e-5 against the oracle in float32."""

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


def _smooth(n):
    for r in (2, 3, 5, 7):
        while n % r == 0:
            n //= r
    return n == 1


def _tol(n):
    return 2e-6 if _smooth(n) and n <= 4096 else 1e-5


@pytest.mark.parametrize("n", [2, 3, 7, 30, 97, 1000, 4099, 5000, 8192, 12289, 65536])
def test_general_ffts_1d(dev, n):
    """R2C / C2R (plan forward / inverse) and C2C (FourierMesh.fft / ifft) of 1D lines."""
    import torchfsm_amd as tfa
    from torchfsm_amd.engine import SpectralPlan

    rng = np.random.default_rng(n)
    u = rng.standard_normal((3, 1, n)).astype(np.float32)
    plan = SpectralPlan([(0, 1, n)], 3, 1, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    ref = np.fft.rfft(u.astype(np.float64), axis=-1)
    e_f = _rel(uh.cpu().numpy()[..., : n // 2 + 1], ref)
    e_b = _rel(plan.inverse(uh).cpu().numpy(), u)
    mesh = tfa.FourierMesh([(0, 1, n)], device=dev)
    z = (rng.standard_normal((2, 1, n)) + 1j * rng.standard_normal((2, 1, n))).astype(np.complex64)
    zf = mesh.fft(torch.from_numpy(z).to(dev))
    e_c = _rel(zf.cpu().numpy(), np.fft.fft(z.astype(np.complex128), axis=-1))
    e_i = _rel(mesh.ifft(zf).cpu().numpy(), z)
    print(f"n={n} general={plan.general}: r2c {e_f:.2e} c2r {e_b:.2e} c2c {e_c:.2e} ic2c {e_i:.2e}")
    assert max(e_f, e_b, e_c, e_i) <= _tol(n)


@pytest.mark.parametrize("shape", [(30, 42), (127, 64), (1000, 6), (1024, 1024), (24, 20, 18), (31, 17, 9),
                                   (256, 256, 256)])
def test_general_ffts_2d_3d(dev, shape):
    import torchfsm_amd as tfa
    from torchfsm_amd.engine import SpectralPlan

    d = len(shape)
    rng = np.random.default_rng(sum(shape))
    u = rng.standard_normal((1, 2) + shape).astype(np.float32)
    axes = tuple(range(-d, 0))
    plan = SpectralPlan([(0, 1, n) for n in shape], 1, 2, device=dev)
    uh = plan.forward(torch.from_numpy(u).to(dev))
    ref = np.fft.rfftn(u.astype(np.float64), axes=axes)
    e_f = _rel(uh.cpu().numpy()[..., : shape[-1] // 2 + 1], ref)
    e_b = _rel(plan.inverse(uh).cpu().numpy(), u)
    full = plan.half_to_full(uh).cpu().numpy()
    e_h = _rel(full, np.fft.fftn(u.astype(np.float64), axes=axes))
    mesh = tfa.FourierMesh([(0, 1, n) for n in shape], device=dev)
    z = (u[:, :1] + 1j * u[:, 1:]).astype(np.complex64)
    e_c = _rel(mesh.fft(torch.from_numpy(z).to(dev)).cpu().numpy(), np.fft.fftn(z.astype(np.complex128), axes=axes))
    tol = max(_tol(n) for n in shape)
    print(f"{shape}: r2c {e_f:.2e} c2r {e_b:.2e} half->full {e_h:.2e} c2c {e_c:.2e}")
    assert max(e_f, e_b, e_h, e_c) <= tol


CASES = {
    # name: (kind, parameter, mesh_info, channels, batch, steps, dt, scheme)
    "burgers1d_5000_fourstep": ("burgers", 0.01, [(0, 1, 5000)], 1, 2, 10, 1e-4, "SETDRK4"),
    "burgers1d_97_bluestein": ("burgers", 0.01, [(0, 1, 97)], 1, 2, 20, 1e-3, "ETDRK2"),
    "kshd1d_300": ("kshd", None, [(0, 32 * math.pi, 300)], 1, 3, 20, 0.05, "SETDRK4"),
    "burgers2d_60x72": ("burgers", 0.01, [(0, 1, 60), (0, 1, 72)], 2, 2, 10, 2e-3, "SETDRK2"),
    "kshd2d_50": ("kshd", None, [(0, 20 * math.pi, 50)] * 2, 1, 2, 10, 0.02, "SETDRK4"),
    "nsp2d_40": ("nsp", 100.0, [(0, 2 * math.pi, 40)] * 2, 2, 2, 10, 1e-2, "SETDRK4"),
    "burgers3d_20x24x30": ("burgers", 0.01, [(0, 1, 20), (0, 1, 24), (0, 1, 30)], 3, 2, 5, 2e-3, "SETDRK4"),
    "nsp3d_24": ("nsp", 100.0, [(0, 2 * math.pi, 24)] * 3, 3, 1, 5, 1e-2, "SETDRK2"),
    "cons3d_32": ("cons", 0.02, [(0, 1, 32)] * 3, 3, 2, 5, 2e-3, "SETDRK4"),
}


def _case_op(kind, par):
    from torchfsm_amd import pde
    from torchfsm_amd.operator import ConservativeConvection, Laplacian
    from oracle import refpath

    if kind == "burgers":
        return pde.Burgers(par), refpath.burgers_terms(par)
    if kind == "kshd":
        return pde.KuramotoSivashinskyHighDim(), refpath.ks_highdim_terms()
    if kind == "nsp":
        return pde.NavierStokes(par), refpath.ns_velocity_terms(par)
    return par * Laplacian() - ConservativeConvection(), [(par, "lap", None), (-1, "cons", None)]


@pytest.mark.parametrize("name", sorted(CASES))
def test_general_stepping_vs_oracle(dev, name):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.integrator import ETDRKIntegrator, SETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    kind, par, mi, c, b, steps, dt, scheme = CASES[name]
    op, terms = _case_op(kind, par)
    op.set_integrator(getattr(ETDRKIntegrator, scheme, None) or getattr(SETDRKIntegrator, scheme))
    dims = tuple(int(m[2]) for m in mi)
    u = 0.5 * fourier_modes_field(dims if len(dims) > 1 else dims[0], len(dims), c, b, seed=100 + sorted(CASES).index(name))
    mesh = tfa.MeshGrid(mi, device=dev)
    out = op.integrate(u.to(dev), mesh=mesh, dt=dt, step=steps)
    ref = refpath.integrate(terms, mi, u.to(dev), dt, steps, scheme=scheme)
    plan = op._prepare(b, dev, dt)
    e = _rel(out, ref)
    print(f"{name}: general={plan.general} rel-L2 vs oracle {e:.2e}")
    assert plan.general
    assert e <= TOL


def test_ns2d_1024_vs_oracle(dev):
    """NS vorticity with Kolmogorov forcing on 1024^2 (B=2, SETDRK4 'auto', dt 5e-4, 10 steps)."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import field, pde
    from torchfsm_amd.workloads import ns_initial_vorticity

    n = 1024
    mesh = tfa.MeshGrid([(0, 2 * math.pi, n)] * 2, device=dev)
    y = mesh.bc_mesh_grid()[1]
    op = pde.NavierStokesVorticity(1000, field.kolm_force(y))
    w0 = ns_initial_vorticity(n, 2, seed=7)
    out = op.integrate(w0.to(dev), mesh=mesh, dt=5e-4, step=10)
    ref = refpath.integrate(refpath.ns_terms(1000, y), [(0, 2 * math.pi, n)] * 2, w0.to(dev), 5e-4, 10,
                            scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"NS 1024^2: general={op._prepare(2, dev, 5e-4).general} rel-L2 vs oracle {e:.2e}")
    assert e <= TOL


def test_burgers3d_256_vs_oracle(dev):
    """3D vector Burgers on 256^3 (B=1, ETDRK2, dt 5e-3, 3 steps)."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    n = 256
    mi = [(0, 1, n)] * 3
    u = fourier_modes_field(n, 3, 3, 1, seed=3)
    op = pde.Burgers(0.01)
    op.set_integrator(ETDRKIntegrator.ETDRK2)
    out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=5e-3, step=3)
    ref = refpath.integrate(refpath.burgers_terms(0.01), mi, u.to(dev), 5e-3, 3, scheme="ETDRK2")
    e = _rel(out, ref)
    print(f"Burgers 256^3: rel-L2 vs oracle {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("dims", [2, 3])
def test_complex_l_with_product_term(dev, dims):
    """A first-derivative (imaginary symbol) term next to a product term: advection by a uniform
    flow added to NS vorticity in 2D and to KS (KSConvection) in 3D."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.operator import SpatialDerivative
    from torchfsm_amd.workloads import fourier_modes_field

    if dims == 2:
        # low modes only: at the Nyquist row of the derivative axis the reference's state is not
        # Hermitian under an imaginary symbol, which a half spectrum does not hold
        n, mi = 64, [(0, 2 * math.pi, 64)] * 2
        op = pde.NavierStokesVorticity(500) - 0.7 * SpatialDerivative(0, 1)
        terms = [(-1, "vc", None), (1 / 500, "lap", None), (-0.7, "deriv", (0, 1))]
        u, dt, steps = fourier_modes_field(n, 2, 1, 2, seed=8), 2e-3, 10
    else:
        n, mi = 16, [(0, 12 * math.pi, 16)] * 3
        op = pde.KuramotoSivashinskyHighDim() + 0.5 * SpatialDerivative(2, 1)
        terms = refpath.ks_highdim_terms() + [(0.5, "deriv", (2, 1))]
        u, dt, steps = 0.3 * fourier_modes_field(n, 3, 1, 2, seed=9), 0.02, 10
    out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=dt, step=steps)
    ref = refpath.integrate(terms, mi, u.to(dev), dt, steps, scheme="SETDRK4")
    e = _rel(out, ref)
    print(f"complex L + product {dims}D: rel-L2 vs oracle {e:.2e}")
    assert e <= TOL


def test_forced_unfused_path_matches_fused(dev):
    """The unfused path on meshes the fused kernels cover: NS 64^2 (SETDRK4), Burgers 3D 32^3
    (ETDRK2) and KS 1D 256 (SETDRK4) agree with the fused kernels and with the oracle."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import field, pde
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field, ns_initial_vorticity

    mesh = tfa.MeshGrid([(0, 2 * math.pi, 64)] * 2, device=dev)
    y = mesh.bc_mesh_grid()[1]
    ns = pde.NavierStokesVorticity(1000, field.kolm_force(y))
    cases = [
        (ns, ns_initial_vorticity(64, 2, seed=11), [(0, 2 * math.pi, 64)] * 2, refpath.ns_terms(1000, y.cpu()), 1e-3,
         20, "SETDRK4"),
        (pde.Burgers(0.01), fourier_modes_field(32, 3, 3, 2, seed=12), [(0, 1, 32)] * 3, refpath.burgers_terms(0.01),
         5e-3, 10, "ETDRK2"),
        (pde.KuramotoSivashinsky(), fourier_modes_field(256, 1, 1, 2, seed=13), [(0, 32 * math.pi, 256)],
         refpath.ks_terms(), 0.05, 20, "SETDRK4"),
    ]
    for op, u, mi, terms, dt, steps, scheme in cases:
        if scheme == "ETDRK2":
            op.set_integrator(ETDRKIntegrator.ETDRK2)
        op.register_mesh(tfa.FourierMesh(mi, device=dev), u.shape[1])
        plan = op._prepare(u.shape[0], dev, dt)
        assert not plan.general
        uh = plan.forward(u.to(dev))
        plan.step(uh, steps)
        fused = plan.inverse(uh)
        plan.set_general(True)
        assert plan.general and not plan.persistent
        uh = plan.forward(u.to(dev))
        plan.step(uh, steps)
        unfused = plan.inverse(uh)
        plan.set_general(False)
        ref = refpath.integrate(terms, mi, u, dt, steps, scheme=scheme)
        e_fu, e_u, e_f = _rel(unfused, fused), _rel(unfused, ref), _rel(fused, ref)
        print(f"{len(mi)}D {mi[0][2]}: unfused vs fused {e_fu:.2e}; vs oracle unfused {e_u:.2e} fused {e_f:.2e}")
        assert e_fu <= TOL and e_u <= TOL


def test_unfused_graph_replay_and_recorder(dev):
    """A general-size plan (NS 48x60, SETDRK4) replays a captured step graph and records a
    trajectory: the same frames as the oracle."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import pde
    from torchfsm_amd.traj_recorder import AutoRecorder, IntervalController
    from torchfsm_amd.workloads import fourier_modes_field

    mi = [(0, 2 * math.pi, 48), (0, 2 * math.pi, 60)]
    w0 = fourier_modes_field((48, 60), 2, 1, 2, seed=14)
    op = pde.NavierStokesVorticity(200)
    rec = AutoRecorder(IntervalController(4))
    traj = op.integrate(w0.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=2e-3, step=40, trajectory_recorder=rec)
    ref = refpath.integrate([(-1, "vc", None), (1 / 200, "lap", None)], mi, w0, 2e-3, 40, scheme="SETDRK4",
                            record=lambda s: s % 4 == 0)
    plan = op._prepare(2, dev, 2e-3)
    e = _rel(traj, ref)
    print(f"NS 48x60 general={plan.general} recorder rel-L2 {e:.2e}")
    assert plan.general and e <= TOL
    out = op.integrate(w0.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=2e-3, step=40)
    plan = op._prepare(2, dev, 2e-3)  # a mesh argument registers the mesh again: a new plan
    assert getattr(plan, "_step_graph", None) is not None
    e2 = _rel(out, ref[:, -1])
    print(f"NS 48x60 graph replay rel-L2 {e2:.2e}")
    assert e2 <= TOL


@pytest.mark.parametrize("dims", [2, 3])
def test_forced_ns_velocity_vs_reference_golden(dev, dims):
    """NavierStokes(100, force=ExplicitSource(f)) (pde.py:85-100) against the reference's own runs
    (golden g10): every ETD scheme in float32, SETDRK4 against its float64 run, and __call__.  The
    forced core masks the integrator's stage tensors in place and adds the force twice
    (_navier_stokes.py:238-262); u0 has content at every frequency, so the masking matters."""
    import torchfsm_amd as tfa
    from conftest import golden
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator, SETDRKIntegrator
    from torchfsm_amd.operator import ExplicitSource

    g = golden("g10_nsp_force.npz")
    tag = f"{dims}d"
    u0 = torch.from_numpy(g[f"u0_{tag}"]).float().to(dev)
    f = torch.from_numpy(g[f"force_{tag}"]).float().to(dev)
    n = u0.shape[-1]
    mesh = tfa.MeshGrid([(0, 1, n)] * dims, device=dev)
    for scheme in ("ETDRK1", "ETDRK2", "SETDRK1", "SETDRK2", "SETDRK3", "SETDRK4"):
        op = pde.NavierStokes(100, force=ExplicitSource(f))
        op.set_integrator(getattr(ETDRKIntegrator, scheme, None) or getattr(SETDRKIntegrator, scheme))
        out = op.integrate(u0, mesh=mesh, dt=2e-3, step=6)
        assert op._prepare(u0.shape[0], dev, 2e-3).general
        e = _rel(out, g[f"out_{tag}_{scheme.lower()}_f32"])
        msg = f"forced NS {dims}D {scheme}: rel-L2 vs reference f32 {e:.2e}"
        if scheme == "SETDRK4":
            e64 = _rel(out, g[f"out_{tag}_setdrk4_f64"])
            e_ref = _rel(g[f"out_{tag}_setdrk4_f32"], g[f"out_{tag}_setdrk4_f64"])
            msg += f"; vs f64 {e64:.2e} (reference f32 vs f64 {e_ref:.2e})"
            assert e64 <= TOL
        print(msg)
        assert e <= TOL
    op = pde.NavierStokes(100, force=ExplicitSource(f))
    call = op(u0, mesh=mesh)
    ref = np.fft.ifftn(g[f"call_{tag}_f64"], axes=tuple(range(-dims, 0))).real
    e = _rel(call, ref)
    print(f"forced NS {dims}D __call__: rel-L2 vs reference f64 {e:.2e}")
    assert e <= TOL


def test_fused_1024_columns_and_rows(dev):
    """2D meshes with 1024-point axes run the fused kernels (8-column tiles, 8 values per thread):
    NS vorticity 1024^2 (B=2, SETDRK4) against the unfused path and the oracle, and 2-channel
    Burgers on 1024 x 256 and 256 x 1024 (ETDRK2) against the oracle."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import field, pde
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field, ns_initial_vorticity

    n = 1024
    mi = [(0, 2 * math.pi, n)] * 2
    mesh = tfa.MeshGrid(mi, device=dev)
    y = mesh.bc_mesh_grid()[1]
    op = pde.NavierStokesVorticity(1000, field.kolm_force(y))
    op.register_mesh(tfa.FourierMesh(mesh), 1)
    plan = op._prepare(2, dev, 5e-4)
    assert not plan.general and plan.tuning[:2] == (8, 1)
    w0 = ns_initial_vorticity(n, 2, seed=17).to(dev)
    uh = plan.forward(w0)
    plan.step(uh, 10)
    fused = plan.inverse(uh)
    plan.set_general(True)
    uh = plan.forward(w0)
    plan.step(uh, 10)
    unfused = plan.inverse(uh)
    plan.set_general(False)
    ref = refpath.integrate(refpath.ns_terms(1000, y), mi, w0, 5e-4, 10, scheme="SETDRK4")
    e_f, e_u, e_fu = _rel(fused, ref), _rel(unfused, ref), _rel(fused, unfused)
    print(f"NS 1024^2: fused vs oracle {e_f:.2e}, unfused vs oracle {e_u:.2e}, fused vs unfused {e_fu:.2e}")
    assert max(e_f, e_u, e_fu) <= TOL
    for shape in ((1024, 256), (256, 1024)):
        mi = [(0, 1, shape[0]), (0, 1, shape[1])]
        u = 0.5 * fourier_modes_field(shape, 2, 2, 1, seed=18)
        op = pde.Burgers(0.01)
        op.set_integrator(ETDRKIntegrator.ETDRK2)
        out = op.integrate(u.to(dev), mesh=tfa.MeshGrid(mi, device=dev), dt=1e-3, step=10)
        assert not op._prepare(1, dev, 1e-3).general
        e = _rel(out, refpath.integrate(refpath.burgers_terms(0.01), mi, u, 1e-3, 10, scheme="ETDRK2"))
        print(f"Burgers {shape}: fused vs oracle {e:.2e}")
        assert e <= TOL
