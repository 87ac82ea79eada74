"""The diagnostic operators (Div, Grad, Curl, Vorticity2Velocity, Vorticity2Pressure,
Velocity2Pressure, ConservativeConvection) and every function of the functional module
(functional.py:9-360) on the engine against the oracle on the same float32 inputs.  A single
derivative multiplies the float32 round-off of the input's top modes by |a|, so each result must be
within rel-L2 1e-5 of the reference's float32 value or at most twice the reference's own float32
error against float64."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a).astype(np.complex128)
    b = np.asarray(b.detach().cpu() if isinstance(b, torch.Tensor) else b).astype(np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check(name, got, f32, f64):
    e, er, e64 = _rel(got, f32), _rel(f32, f64), _rel(got, f64)
    print(f"{name}: rel-L2 vs reference f32 {e:.2e}; vs f64 {e64:.2e} (reference f32 vs f64 {er:.2e})")
    assert got.shape == f32.shape, (got.shape, f32.shape)
    assert e <= 1e-5 or e64 <= 2 * er + 1e-6


def _field(n, dims, ch, batch, seed):
    from torchfsm_amd.workloads import fourier_modes_field

    return fourier_modes_field(n, dims, ch, batch, seed=seed)


def _ref(fn, u, mi, *args):
    """fn(u^, mesh) in float32 and float64, inverse transformed."""
    from oracle import refpath

    out = []
    for dt in (torch.float32, torch.float64):
        m = refpath.Mesh(mi, dtype=dt)
        out.append(m.ifft(fn(m.fft(u.to(dt)), m, *args)).real)
    return out


CASES = [(1, 128), (2, 64), (3, 32)]


@pytest.mark.parametrize("dims,n", CASES)
def test_div_grad_curl(dev, dims, n):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import functional as F

    mi = [(0, 2 * math.pi, n)] * dims
    mesh = tfa.MeshGrid(mi, device=dev)
    v = _field(n, dims, dims, 2, seed=dims)
    s = _field(n, dims, 1, 2, seed=10 + dims)
    _check(f"div {dims}D", F.div(v.to(dev), mesh=mesh), *_ref(refpath.div, v, mi))
    _check(f"grad {dims}D", F.grad(s.to(dev), mesh=mesh), *_ref(refpath.grad, s, mi))
    if dims >= 2:
        _check(f"curl {dims}D", F.curl(v.to(dev), mesh=mesh), *_ref(refpath.curl, v, mi))


def test_vorticity_velocity_pressure_2d(dev):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import functional as F
    from torchfsm_amd.operator import Vorticity2Pressure, Vorticity2Velocity

    n = 64
    mi = [(0, 2 * math.pi, n)] * 2
    mesh = tfa.MeshGrid(mi, device=dev)
    w = _field(n, 2, 1, 2, seed=21)
    v2v = lambda f, m: f * refpath.vorticity2velocity_coef(m)  # noqa: E731
    _check("vorticity2velocity", F.vorticity2velocity(w.to(dev), mesh=mesh), *_ref(v2v, w, mi))
    _check("Vorticity2Velocity", Vorticity2Velocity()(w.to(dev), mesh=mesh), *_ref(v2v, w, mi))
    _check("vorticity2pressure", F.vorticity2pressure(w.to(dev), mesh=mesh), *_ref(refpath.vorticity2pressure, w, mi))
    _check("Vorticity2Pressure", Vorticity2Pressure()(w.to(dev), mesh=mesh), *_ref(refpath.vorticity2pressure, w, mi))


@pytest.mark.parametrize("dims,n", CASES)
def test_velocity2pressure_and_conservative_convection(dev, dims, n):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import functional as F

    mi = [(0, 2 * math.pi, n)] * dims
    mesh = tfa.MeshGrid(mi, device=dev)
    v = _field(n, dims, dims, 2, seed=30 + dims)
    _check(f"velocity2pressure {dims}D", F.velocity2pressure(v.to(dev), mesh=mesh),
           *_ref(refpath.velocity2pressure, v, mi))
    _check(f"conservative_convection {dims}D", F.conservative_convection(v.to(dev), mesh=mesh),
           *_ref(refpath.conservative_convection, v, mi))


def test_functional_wrappers_of_engine_operators(dev):
    """laplacian, biharmonic, spatial_derivative, convection, vorticity_convection, ks_convection."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import functional as F

    n = 64
    mi = [(0, 2 * math.pi, n)] * 2
    mesh = tfa.MeshGrid(mi, device=dev)
    s = _field(n, 2, 1, 2, seed=41)
    v = _field(n, 2, 2, 2, seed=42)
    cases = [("laplacian", F.laplacian(s.to(dev), mesh=mesh), [(1, "lap", None)], s),
             ("biharmonic", F.biharmonic(s.to(dev), mesh=mesh), [(1, "biharm", None)], s),
             ("spatial_derivative", F.spatial_derivative(1, 3, s.to(dev), mesh=mesh), [(1, "deriv", (1, 3))], s),
             ("convection", F.convection(v.to(dev), mesh=mesh), [(1, "conv", None)], v),
             ("vorticity_convection", F.vorticity_convection(s.to(dev), mesh=mesh), [(1, "vc", None)], s),
             ("ks_convection", F.ks_convection(s.to(dev), mesh=mesh), [(1, "ks", None)], s)]
    for name, got, terms, x in cases:
        _check(name, got, refpath.apply(terms, mi, x), refpath.apply(terms, mi, x.double()))


def test_fourier_io_coefficients_and_errors(dev):
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.operator import Div, Grad, Laplacian

    n = 64
    mi = [(0, 2 * math.pi, n)] * 2
    mesh = tfa.MeshGrid(mi, device=dev)
    v = _field(n, 2, 2, 2, seed=51)
    m = refpath.Mesh(mi)
    vf = m.fft(v)
    got = (2.5 * Div())(u_fft=vf.to(dev), mesh=mesh, return_in_fourier=True)
    ref = 2.5 * refpath.div(vf, m)
    print(f"2.5 * Div, Fourier in/out: rel-L2 {_rel(got, ref):.2e}")
    assert got.shape == ref.shape and _rel(got, ref) <= 1e-5
    with pytest.raises(ValueError):
        (Div() + Laplacian())(v.to(dev), mesh=mesh)
    with pytest.raises(ValueError):
        Grad().integrate(torch.zeros(1, 1, n, n, device=dev), mesh=mesh, dt=0.1, step=1)
    with pytest.raises(ValueError, match="dimension"):
        Div()(torch.zeros(1, 1, n, n, device=dev), mesh=mesh)


@pytest.mark.parametrize("dims,n", [(1, 256), (1, 100), (2, 64), (2, 100), (2, 512)])
def test_conservative_burgers_integrates(dev, dims, n):
    """nu Laplacian - ConservativeConvection stepped by the engine (SETDRK4 under every launch shape
    the autotuner tries, and ETDRK2) against the oracle, and its Operator.__call__."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd import autotune
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.operator import ConservativeConvection, Laplacian

    u0 = 0.5 * _field(n, dims, dims, 2, seed=60 + n)
    mi = [(0, 1, n)] * dims
    mesh = tfa.MeshGrid(mi, device=dev)
    terms = [(0.01, "lap", None), (-1, "cons", None)]
    dt, steps = (2e-4, 20) if dims == 1 else (5e-4, 10)
    ref = refpath.integrate(terms, mi, u0, dt, steps, scheme="SETDRK4")
    op = 0.01 * Laplacian() - ConservativeConvection()
    e = _rel(op.integrate(u0.to(dev), mesh=mesh, dt=dt, step=steps), ref)
    print(f"conservative Burgers {dims}D {n} SETDRK4: rel-L2 {e:.2e}")
    assert e <= 1e-5
    if dims == 2:
        plan = op._prepare(2, dev, dt)
        for cand in autotune.candidates(plan):
            plan.set_tuning(*cand)
            uh = plan.forward(u0.to(dev))
            plan.step(uh, steps)
            e1 = _rel(plan.inverse(uh), ref)
            print(f"  launch shape {cand}: rel-L2 {e1:.2e}")
            assert e1 <= 1e-5
    op.set_integrator(ETDRKIntegrator.ETDRK2)
    e2 = _rel(op.integrate(u0.to(dev), mesh=mesh, dt=dt, step=steps),
              refpath.integrate(terms, mi, u0, dt, steps, scheme="ETDRK2"))
    print(f"conservative Burgers {dims}D {n} ETDRK2: rel-L2 {e2:.2e}")
    assert e2 <= 1e-5
    _check(f"conservative Burgers call {dims}D {n}", op(u0.to(dev), mesh=mesh), refpath.apply(terms, mi, u0),
           refpath.apply(terms, mi, u0.double()))


def test_conservative_convection_3d_steps_on_the_unfused_path(dev):
    """nu Laplacian - ConservativeConvection in 3D (16^3, SETDRK4 and ETDRK2) against the oracle: the
    fused 3D kernels do not carry this term, the unfused path does."""
    import torchfsm_amd as tfa
    from oracle import refpath
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.operator import ConservativeConvection, Laplacian

    n, dt, steps = 16, 5e-4, 10
    u0 = 0.5 * _field(n, 3, 3, 2, seed=77)
    mi = [(0, 1, n)] * 3
    mesh = tfa.MeshGrid(mi, device=dev)
    terms = [(0.01, "lap", None), (-1, "cons", None)]
    op = 0.01 * Laplacian() - ConservativeConvection()
    for scheme in ("SETDRK4", "ETDRK2"):
        if scheme == "ETDRK2":
            op.set_integrator(ETDRKIntegrator.ETDRK2)
        out = op.integrate(u0.to(dev), mesh=mesh, dt=dt, step=steps)
        assert op._prepare(2, dev, dt).general
        e = _rel(out, refpath.integrate(terms, mi, u0, dt, steps, scheme=scheme))
        print(f"conservative Burgers 3D {n} {scheme}: rel-L2 {e:.2e}")
        assert e <= 1e-5
