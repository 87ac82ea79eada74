"""The oracle, pinned: refpath against the reference's own golden vectors (tests/golden, written by
importing the reference), numref against refpath in float64, and numref against closed forms,
brute force, invariants and convergence orders.  CPU only."""

import math

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import numref
from oracle import refpath as R


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def kolm_y(n, dtype):
    return (2 * math.pi * torch.arange(0, n, dtype=dtype) / n).reshape(1, 1, 1, n).repeat(1, 1, n, 1)


# ---------------------------------------------------------------- refpath == reference (golden)
@pytest.mark.parametrize("name,dt", [("f32", torch.float32), ("f64", torch.float64)])
def test_refpath_burgers1d_trajectory(name, dt):
    g = golden("g1_burgers1d.npz")
    u0 = torch.from_numpy(g["u0_" + name])
    t = R.integrate(R.burgers_terms(0.01), [(0, 1, 128)], u0, 0.01, 200, record=lambda s: True)
    assert t.shape == (1, 201, 1, 128)
    assert rel(t.numpy(), g["traj_" + name]) < 1e-6


@pytest.mark.parametrize("name,dt", [("f32", torch.float32), ("f64", torch.float64)])
def test_refpath_ns64(name, dt):
    g = golden("g2_ns64.npz")
    o = R.integrate(R.ns_terms(1000, kolm_y(64, dt)), [(0, 2 * math.pi, 64)] * 2, torch.from_numpy(g["u0_" + name]),
                    1e-3, 20, scheme="SETDRK4")
    assert rel(o.numpy(), g["out_" + name]) < 1e-6


def test_refpath_ns128_100_steps():
    g = golden("g9_ns128.npz")
    o = R.integrate(R.ns_terms(1000, kolm_y(128, torch.float32)), [(0, 2 * math.pi, 128)] * 2,
                    torch.from_numpy(g["u0_f32"]), 1e-3, 100)
    assert rel(o.numpy(), g["out_f32"]) < 1e-6


@pytest.mark.parametrize("scheme", ["etdrk2", "setdrk2"])
def test_refpath_burgers3d(scheme):
    g = golden("g3_burgers3d.npz")
    for name in ("f32", "f64"):
        o = R.integrate(R.burgers_terms(0.01), [(0, 1, 16)] * 3, torch.from_numpy(g["u0_" + name]), 5e-3, 10,
                        scheme=scheme.upper())
        assert rel(o.numpy(), g[f"out_{scheme}_{name}"]) < 1e-6


def test_refpath_burgers2d_autograd():
    g = golden("g4_burgers2d_grad.npz")
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        nu = torch.tensor(0.01, dtype=dt, requires_grad=True)
        u = torch.from_numpy(g["u0_" + name]).clone().requires_grad_(True)
        o = R.integrate(R.burgers_terms(nu), [(0, 1, 32)] * 2, u, 0.01, 10, scheme="SETDRK2")
        loss = ((o - torch.from_numpy(g["target_" + name])) ** 2).mean()
        loss.backward()
        assert rel(o.detach().numpy(), g["out_" + name]) < 1e-6
        assert rel(u.grad.numpy(), g["grad_u0_" + name]) < 1e-6
        assert abs(float(nu.grad) - float(g["grad_nu_" + name])) < 1e-5 * abs(float(g["grad_nu_" + name]))


def test_refpath_coefficients():
    g = golden("g5_coefs.npz")
    mesh = R.Mesh([(0, 2 * math.pi, 64)] * 2)
    terms = R.ns_terms(1000, kolm_y(64, torch.float32))
    L = R.build_linear(mesh, terms, 1)
    assert np.array_equal(L.numpy(), g["ns_L"])
    sch = R.Scheme("SETDRK4", 1e-3, L, None)
    for k, key in (("E", "ns_E"), ("E2", "ns_E2"), ("c1", "ns_c1"), ("c4", "ns_c4"), ("c5", "ns_c5"), ("c6", "ns_c6")):
        assert rel(getattr(sch, k).numpy(), g[key]) < 1e-6
    mesh = R.Mesh([(0, 1, 128)] * 2)
    L = R.build_linear(mesh, R.burgers_terms(0.01), 2)
    sch = R.Scheme("ETDRK2", 0.01, L, None)
    assert np.array_equal(sch.c1.numpy(), g["b2_c1"]) and np.array_equal(sch.c2.numpy(), g["b2_c2"])


def test_refpath_nonlinear_terms():
    g = golden("g6_nonlinear.npz")
    o = R.apply([(1, "vc", None)], [(0, 2 * math.pi, 64)] * 2, torch.from_numpy(g["vc_in"]), return_in_fourier=True)
    assert rel(o.numpy(), g["vc_out"]) < 1e-6
    for tag, n, d in (("c1", 64, 1), ("c2", 32, 2), ("c3", 16, 3)):
        o = R.apply([(1, "conv", None)], [(0, 1, n)] * d, torch.from_numpy(g[tag + "_in"]), return_in_fourier=True)
        assert rel(o.numpy(), g[tag + "_out"]) < 1e-6


def test_refpath_ks_and_linear():
    g = golden("g7_ks1d.npz")
    o = R.integrate(R.ks_terms(), [(0, 32 * math.pi, 512)], torch.from_numpy(g["u0"]), 0.1, 10, scheme="SETDRK4")
    assert rel(o.numpy(), g["out"]) < 1e-6
    g = golden("g8_linear.npz")
    o = R.integrate([(0.01, "lap", None)], [(0, 1, 32)] * 2, torch.from_numpy(g["u0"]), 0.05, 5)
    assert rel(o.numpy(), g["out"]) < 1e-6


# ---------------------------------------------------------------- numref == refpath in float64
def test_numref_matches_refpath_f64():
    g = golden("g1_burgers1d.npz")
    t = numref.integrate([(0, 1, 128)], [(0.01, "lap", None), (-1, "conv", None)], g["u0_f64"], 0.01, 200,
                         record_every=1)
    assert rel(t, g["traj_f64"]) < 1e-12
    g = golden("g2_ns64.npz")
    y = np.broadcast_to(2 * np.pi * np.arange(64) / 64, (1, 1, 64, 64))
    terms = [(-1, "vc", None), (1 / 1000, "lap", None), (-0.1, "unit", None), (-1, "src", 4 * np.cos(4 * y))]
    o = numref.integrate([(0, 2 * np.pi, 64)] * 2, terms, g["u0_f64"], 1e-3, 20, "SETDRK4")
    assert rel(o, g["out_f64"]) < 1e-12
    g = golden("g3_burgers3d.npz")
    for s in ("etdrk2", "setdrk2"):
        o = numref.integrate([(0, 1, 16)] * 3, [(0.01, "lap", None), (-1, "conv", None)], g["u0_f64"], 5e-3, 10,
                             s.upper())
        assert rel(o, g[f"out_{s}_f64"]) < 1e-11


# ---------------------------------------------------------------- closed forms, brute force, invariants
def test_fft_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(32) + 1j * rng.standard_normal(32)
    assert np.abs(numref.dft_brute(x) - np.fft.fft(x)).max() < 1e-12
    # torch.fft (refpath's transform) against the brute-force DFT
    assert np.abs(torch.fft.fft(torch.from_numpy(x)).numpy() - numref.dft_brute(x)).max() < 1e-12


def test_phi_functions_closed_forms():
    z = np.array([1e-8, -1e-3, 0.3 + 0.2j, -0.49, -0.51, 2.0, -30.0, 5j])
    for order, ref in ((1, lambda z: (np.exp(z) - 1) / z), (2, lambda z: (np.exp(z) - 1 - z) / z**2),
                       (3, lambda z: (np.exp(z) - 1 - z - z**2 / 2) / z**3)):
        got = numref.phi(z, order)
        big = np.abs(z) > 1e-2
        assert np.allclose(got[big], ref(z[big]), rtol=1e-12, atol=0)
        assert abs(got[0] - 1 / math.factorial(order)) < 1e-8
    # the contour mean of the reference (16 points, radius 1) equals the phi forms in float64
    L = torch.tensor([0.0, -1e-3, -0.5, -3.0, -40.0, 0.2], dtype=torch.float64)
    sch = R.Scheme("SETDRK4", 1.0, L, None)
    ex = numref.coefficients("SETDRK4", 1.0, L.numpy())
    for k in ("c1", "c4", "c5", "c6"):
        assert np.allclose(getattr(sch, k).numpy(), ex[k], rtol=1e-10, atol=1e-13)


def test_dealias_mask_table():
    for n, kept in ((64, 21), (128, 42), (256, 85), (512, 170), (1024, 341), (2048, 682), (4096, 1365)):
        m = R.Mesh([(0, 2 * math.pi, n)]).mask(2 / 3).reshape(-1).numpy()
        k = np.abs(np.fft.fftfreq(n, 1.0 / n))
        assert (m == (k <= kept)).all(), n
        assert np.array_equal(numref.Box([(0, 2 * math.pi, n)]).mask().reshape(-1), (k[: n // 2 + 1] <= kept))


def test_vorticity_convection_single_mode_vanishes():
    n = 32
    x = np.arange(n) * 2 * np.pi / n
    w = np.cos(3 * x)[:, None] * np.sin(2 * x)[None, :]
    box = numref.Box([(0, 2 * np.pi, n)] * 2)
    out = numref.vorticity_convection(box, box.rfft(w[None, None]))
    assert np.abs(out).max() < 1e-10
    o = R.apply([(1, "vc", None)], [(0, 2 * math.pi, n)] * 2, torch.from_numpy(w[None, None]))
    assert o.abs().max() < 1e-10


def test_convection_known_answer():
    """(u.grad)u for u = (sin x, cos y): ((sin x cos x), (-cos y sin y))."""
    n = 32
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    u = np.stack([np.sin(X), np.cos(Y)])[None]
    ref = np.stack([np.sin(X) * np.cos(X), -np.cos(Y) * np.sin(Y)])[None]
    box = numref.Box([(0, 2 * np.pi, n)] * 2)
    got = box.irfft(numref.convection(box, box.rfft(u)))
    assert np.abs(got - ref).max() < 1e-12
    got = R.apply([(1, "conv", None)], [(0, 2 * math.pi, n)] * 2, torch.from_numpy(u)).numpy()
    assert np.abs(got - ref).max() < 1e-12


def test_burgers_mean_conserved():
    g = golden("g1_burgers1d.npz")
    t = g["traj_f64"][0, :, 0, :]
    assert np.abs(t.mean(axis=1) - t[0].mean()).max() < 1e-12


def test_heat_equation_exact():
    """ETDRK0 on nu*lap: every Fourier mode decays exactly as exp(-nu k^2 t)."""
    n, nu, dt, steps = 32, 0.01, 0.05, 7
    x = np.arange(n) / n
    u0 = np.cos(2 * np.pi * 3 * x)[None, None]
    o = numref.integrate([(0, 1, n)], [(nu, "lap", None)], u0, dt, steps)
    ex = u0 * np.exp(-nu * (2 * np.pi * 3) ** 2 * dt * steps)
    assert np.abs(o - ex).max() < 1e-12
    o2 = R.integrate([(nu, "lap", None)], [(0, 1, n)], torch.from_numpy(u0), dt, steps).numpy()
    assert np.abs(o2 - ex).max() < 1e-12


def test_constant_forcing_exact_for_etd1():
    """u' = L u + s with constant s: ETD1 is exact, u(t) = e^{Lt} u0 + t phi1(Lt) s."""
    n, nu, dt = 16, 0.05, 0.1
    x = np.arange(n) / n
    src = np.cos(2 * np.pi * 2 * x)[None, None]
    u0 = np.sin(2 * np.pi * x)[None, None]
    terms = [(nu, "lap", None), (1.0, "src", src)]
    o = numref.integrate([(0, 1, n)], terms, u0, dt, 5, "SETDRK1")
    lam = -nu * (2 * np.pi * 2) ** 2
    t = 5 * dt
    ex = np.exp(-nu * (2 * np.pi) ** 2 * t) * u0 + t * (np.exp(lam * t) - 1) / (lam * t) * src
    assert np.abs(o - ex).max() < 1e-12


@pytest.mark.parametrize("scheme,order", [("SETDRK2", 2), ("SETDRK4", 4), ("SETDRK3", 3)])
def test_convergence_order(scheme, order):
    """Observed order of the stage formulas on 1D viscous Burgers (numref, float64)."""
    n = 64
    x = np.arange(n) / n
    u0 = (np.sin(2 * np.pi * x) + 0.5)[None, None]
    terms = [(0.02, "lap", None), (-1, "conv", None)]
    T = 0.2
    ref = numref.integrate([(0, 1, n)], terms, u0, T / 400, 400, "SETDRK4")
    errs = []
    for m in (10, 20, 40):
        errs.append(np.abs(numref.integrate([(0, 1, n)], terms, u0, T / m, m, scheme) - ref).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert abs(rates[-1] - order) < 0.5, rates


def test_ic_generators_oracle_closed_forms():
    """diffused_noise: before normalisation the spectrum is the draw's spectrum times
    exp(coef * lap) (one exact heat step, dt = 1).  truncated_fourier_series: the spectrum of the
    output is the Hermitian part of the masked coefficients, zero for |f| > threshold."""
    g = torch.Generator().manual_seed(3)
    mi = [(0, 4.0, 32), (0, 2.0, 16)]
    u0 = torch.randn(2, 1, 32, 16, generator=g, dtype=torch.float64)
    raw = R.diffused_noise_from(u0, mi, 0.05, zero_centered=False, unit_variance=False, unit_magnitude=False)
    m = R.Mesh(mi, dtype=torch.float64)
    want = m.ifft(m.fft(u0) * torch.exp(0.05 * m.lap())).real
    assert torch.allclose(raw, want, rtol=0, atol=1e-12)
    out = R.diffused_noise_from(u0, mi, 0.05)
    assert torch.allclose(out.abs().flatten(1).amax(1), torch.ones(2, dtype=torch.float64))
    assert out.mean(dim=(1, 2, 3)).abs().max() < 1e-12
    mag = torch.rand(2, 1, 32, 16, generator=g, dtype=torch.float64) * 2 - 1
    ang = torch.rand(2, 1, 32, 16, generator=g, dtype=torch.float64) * 2 * math.pi
    t = R.truncated_fourier_series_from(mag, ang, mi, 3, zero_centered=False, unit_variance=False,
                                        unit_magnitude=False)
    X = mag * torch.exp(1j * ang)
    fx = torch.fft.fftfreq(32, 4.0 / 32, dtype=torch.float64).abs().reshape(1, 1, -1, 1)
    fy = torch.fft.fftfreq(16, 2.0 / 16, dtype=torch.float64).abs().reshape(1, 1, 1, -1)
    X = X * ((fx <= 3) & (fy <= 3))
    Xm = torch.roll(torch.flip(X, dims=(-2, -1)), shifts=(1, 1), dims=(-2, -1))  # X(-k)
    herm = (X + torch.conj(Xm)) / 2
    assert torch.allclose(torch.fft.fftn(t, dim=(-2, -1)), herm, atol=1e-12)
    assert torch.fft.fftn(t, dim=(-2, -1))[..., ((fx > 3) | (fy > 3)).squeeze(0).squeeze(0)].abs().max() < 1e-12
    t2 = R.truncated_fourier_series_from(mag, ang, mi, 3, unit_magnitude=False, unit_variance=True)
    assert torch.allclose(t2.flatten(1).std(1), torch.ones(2, dtype=torch.float64))


def test_ns_pressure_convection_is_the_leray_projection():
    """NSPressureConvection (_navier_stokes.py:221-262) equals -(I - k k^T / |k|^2) conv for k != 0
    and -conv at k = 0, so its output is divergence free: sum_i k_i N_i = 0."""
    g = torch.Generator().manual_seed(8)
    for mi in ([(0, 1, 16)] * 2, [(0, 1, 8), (0, 2, 8), (0, 1, 8)]):
        D = len(mi)
        m = R.Mesh(mi, dtype=torch.float64)
        u = torch.randn((2, D) + tuple(x[2] for x in mi), generator=g, dtype=torch.float64)
        uf = m.fft(u)
        N = R.ns_pressure_convection(uf, m)
        conv = R.convection(uf, m)
        f = m.bf_vector().to(torch.float64)
        f2 = (f ** 2).sum(1, keepdim=True)
        proj = conv - f * (f * conv).sum(1, keepdim=True) / torch.where(f2 == 0, 1.0, f2)
        assert torch.allclose(N, -proj, atol=1e-10 * float(conv.abs().max()))
        div = (m.nabla_vector(1) * N).sum(1)
        assert float(div.abs().max()) < 1e-9 * float(conv.abs().max())


# ---------------------------------------------------------------- explicit Runge-Kutta family
# RK4_38Rule is second order as the reference writes it: its last row is a_41 = -1, a_42 = 1
# (_rk.py:165-173), where the classical 3/8 rule has a_41 = 1, a_42 = -1; parity keeps it.
RK_ORDERS = {"Euler": 1, "Midpoint": 2, "Heun12": 2, "Ralston12": 2, "BogackiShampine23": 3, "RK4": 4,
             "RK4_38Rule": 2, "Dorpi45": 5, "Fehlberg45": 5, "CashKarp45": 5}


def _pendulum(name, m, T=2.0):
    u = torch.tensor([1.0, 0.5], dtype=torch.float64)
    for _ in range(m):
        u = R.rk_step(lambda v: torch.stack([v[1], -torch.sin(v[0])]), u, T / m, name)
    return u


@pytest.mark.parametrize("name", sorted(RK_ORDERS))
def test_rk_tableau_consistency_and_order(name):
    """Each tableau row satisfies c_i = sum_j a_ij, the weights sum to one, and the step converges at
    the method's order on the pendulum x' = y, y' = -sin x in float64 (reference solution: RK4 at
    4000 steps), so a wrong entry fails."""
    ca, b = R.RK_TABLEAUX[name]
    for row in ca[:len(b) - 1]:  # rows past len(b) - 1 feed stages the update weights never use
        assert abs(row[0] - sum(row[1:])) < 1e-12, row
    assert abs(sum(b) - 1) < 1e-12
    ref = _pendulum("RK4", 4000)
    errs = [float((_pendulum(name, m) - ref).abs().max()) for m in (8, 16, 32)]
    rate = math.log2(errs[1] / errs[2])
    assert abs(rate - RK_ORDERS[name]) < 0.25, (errs, rate)


def test_rk_linear_is_the_stability_polynomial():
    """RK4 on nu*lap (f = L u): each step multiplies mode k by sum_{j<=4} (dt L)^j / j!
    (dt inside the stability interval for every mode of the mesh)."""
    n, nu, dt, steps = 32, 0.01, 0.005, 7
    x = np.arange(n) / n
    u0 = np.cos(2 * np.pi * 3 * x)[None, None]
    z = -nu * (2 * np.pi * 3) ** 2 * dt
    amp = sum(z ** j / math.factorial(j) for j in range(5)) ** steps
    o = R.integrate([(nu, "lap", None)], [(0, 1, n)], torch.from_numpy(u0), dt, steps, scheme="RK4").numpy()
    assert np.abs(o - amp * u0).max() < 1e-12


def test_rk_agrees_with_etd_on_burgers():
    """f = L u + N(u) carries the nonlinear term with the reference's sign: RK4 and SETDRK4 on 1D
    viscous Burgers agree to the schemes' truncation error at a small step (float64)."""
    n = 64
    x = torch.arange(n, dtype=torch.float64) / n
    u0 = (torch.sin(2 * math.pi * x) + 0.5).reshape(1, 1, n)
    terms = R.burgers_terms(0.02)
    d = []
    for dt, m in ((1e-3, 200), (5e-4, 400)):
        a = R.integrate(terms, [(0, 1, n)], u0, dt, m, scheme="RK4")
        b = R.integrate(terms, [(0, 1, n)], u0, dt, m, scheme="SETDRK4")
        d.append(rel(a.numpy(), b.numpy()))
    assert d[0] < 1e-8 and math.log2(d[0] / d[1]) > 3.5, d
    e = R.integrate(terms, [(0, 1, n)], u0, 1e-3, 200, scheme="Euler")
    assert 1e-4 < rel(e.numpy(), b.numpy()) < 1e-2


# ---------------------------------------------------------------- KSConvection
def test_ks_convection_closed_form_and_batch_mean():
    """1/2 |grad phi|^2 for phi_b = s_b (sin x + cos 2y) on [0, 2 pi)^2: 1/2 s^2 (cos^2 x + 4 sin^2 2y)
    = s^2 (5/4 + 1/4 cos 2x - cos 4y); remove_mean subtracts the mean over the whole batch
    (5/4 mean_b s_b^2), so the samples' DC differ from zero in opposite directions."""
    n = 32
    mi = [(0, 2 * math.pi, n)] * 2
    mesh = R.Mesh(mi, dtype=torch.float64)
    x = torch.arange(n, dtype=torch.float64) * 2 * math.pi / n
    X, Y = torch.meshgrid(x, x, indexing="ij")
    s = torch.tensor([1.0, 2.0], dtype=torch.float64).reshape(2, 1, 1, 1)
    phi = s * (torch.sin(X) + torch.cos(2 * Y))
    base = 5 / 4 + 0.25 * torch.cos(2 * X) - torch.cos(4 * Y)
    keep = mesh.ifft(R.ks_convection(mesh.fft(phi), mesh, remove_mean=False)).real
    assert torch.allclose(keep, s**2 * base, atol=1e-12)
    rem = mesh.ifft(R.ks_convection(mesh.fft(phi), mesh, remove_mean=True)).real
    assert torch.allclose(rem, s**2 * base - 5 / 4 * 2.5, atol=1e-12)
    assert abs(float(rem[0].mean()) + 1.875) < 1e-12 and abs(float(rem[1].mean()) - 1.875) < 1e-12


def test_ks_highdim_1d_is_the_derivative_of_ks():
    """d/dx of phi_t = -phi_xx - phi_xxxx - 1/2 phi_x^2 is the KS equation of v = phi_x with
    Convection; the dealias mask and the ETD combines act per mode, so in float64 the two
    integrations agree mode by mode: v^ = i a phi^ at every step (the mean only moves k = 0)."""
    n, L = 64, 22.0
    mi = [(0, L, n)]
    mesh = R.Mesh(mi, dtype=torch.float64)
    x = torch.arange(n, dtype=torch.float64) * L / n
    phi0 = (torch.cos(2 * math.pi * x / L) + 0.3 * torch.sin(4 * math.pi * x / L)).reshape(1, 1, n)
    v0 = mesh.ifft(mesh.nabla_vector(1) * mesh.fft(phi0)).real
    for remove_mean in (True, False):
        ph = R.integrate(R.ks_highdim_terms(remove_mean), mi, phi0, 0.05, 40, scheme="SETDRK4", return_in_fourier=True)
        v = R.integrate(R.ks_terms(), mi, v0, 0.05, 40, scheme="SETDRK4", return_in_fourier=True)
        dv = mesh.nabla_vector(1) * ph
        assert rel(dv.numpy(), v.numpy()) < 1e-10


# ---------------------------------------------------------------- complex L (SpatialDerivative)
def test_airy_equation_exact():
    """u_t = -c u_xxx (SpatialDerivative(0, 3), imaginary symbol): ETDRK0 moves cos(k x) to
    cos(k x + c k^3 t) exactly."""
    n, L, c, dt, steps = 64, 2 * math.pi, 0.3, 0.05, 9
    x = torch.arange(n, dtype=torch.float64) * L / n
    k = 3.0
    u0 = torch.cos(k * x).reshape(1, 1, n)
    o = R.integrate([(-c, "deriv", (0, 3))], [(0, L, n)], u0, dt, steps, scheme="ETDRK0")
    ex = torch.cos(k * x + c * k**3 * dt * steps).reshape(1, 1, n)
    assert (o - ex).abs().max() < 1e-12


def test_kdv_soliton_converges():
    """The reference's default KdV u_t = -u_xxx + 6 u u_x carries u = -(c/2) sech^2(sqrt(c)/2 (x - c t
    - x0)); ETDRK2 (complex c1, c2) converges to it at second order in float64."""
    n, L, cs, T = 256, 40.0, 1.0, 1.0
    x = torch.arange(n, dtype=torch.float64) * L / n

    def sol(t):
        return (-(cs / 2) / torch.cosh(math.sqrt(cs) / 2 * (x - cs * t - 15.0)) ** 2).reshape(1, 1, n)

    errs = []
    for m in (40, 80):
        o = R.integrate(R.kdv_terms(), [(0, L, n)], sol(0.0), T / m, m, scheme="ETDRK2")
        errs.append(float((o - sol(T)).abs().max()))
    assert errs[1] < 1e-4 and 3.5 < errs[0] / errs[1] < 4.6, errs


def test_setdrk_complex_l_takes_the_real_part():
    """With L = i w the reference's SETDRK coefficient is dt Re(phi_1(i w dt)) = dt sin(w dt)/(w dt)
    (the contour mean is taken .real), while E = exp(i w dt) stays complex."""
    dt = 0.1
    w = torch.tensor([0.5, 10.0, 50.0, 300.0], dtype=torch.float64)
    L = 1j * w
    sch = R.Scheme("SETDRK1", dt, L, lambda u: u * 0)
    th = w * dt
    assert torch.allclose(sch.c1, dt * torch.sin(th) / th, rtol=1e-11, atol=0)
    assert torch.allclose(sch.E, torch.exp(1j * th), rtol=1e-14, atol=0)


# ---------------------------------------------------------------- diagnostic operators
def test_diagnostic_operators_closed_forms():
    """Taylor-Green on [0, 2 pi)^2: psi = sin x sin y, w = -lap psi = 2 sin x sin y, velocity
    u = (d_y psi, -d_x psi) = (sin x cos y, -cos x sin y); div u = 0; curl u = w; grad psi = (cos x
    sin y, sin x cos y); (u . grad) u = div(u u^T) = (sin 2x, sin 2y) / 2; pressure
    -lap^-1 div((u . grad) u) = (cos 2x + cos 2y) / 4."""
    n = 32
    mi = [(0, 2 * math.pi, n)] * 2
    m = R.Mesh(mi, dtype=torch.float64)
    x = torch.arange(n, dtype=torch.float64) * 2 * math.pi / n
    X, Y = torch.meshgrid(x, x, indexing="ij")
    psi = (torch.sin(X) * torch.sin(Y))[None, None]
    w = 2 * psi
    vel = torch.stack([torch.sin(X) * torch.cos(Y), -torch.cos(X) * torch.sin(Y)])[None]
    vf = m.fft(vel)
    sp = lambda f: m.ifft(f).real  # noqa: E731
    assert sp(R.div(vf, m)).abs().max() < 1e-12
    assert (sp(R.curl(vf, m)) - w).abs().max() < 1e-12
    g = torch.stack([torch.cos(X) * torch.sin(Y), torch.sin(X) * torch.cos(Y)])[None]
    assert (sp(R.grad(m.fft(psi), m)) - g).abs().max() < 1e-12
    assert (sp(m.fft(w) * R.vorticity2velocity_coef(m)) - vel).abs().max() < 1e-12
    adv = torch.stack([torch.sin(2 * X), torch.sin(2 * Y)])[None] / 2
    assert (sp(R.conservative_convection(vf, m)) - adv).abs().max() < 1e-12
    p = ((torch.cos(2 * X) + torch.cos(2 * Y)) / 4)[None, None]
    assert (sp(R.velocity2pressure(vf, m)) - p).abs().max() < 1e-12
    assert (sp(R.vorticity2pressure(m.fft(w), m)) - p).abs().max() < 1e-12
    # 3D curl of (sin z, sin x, sin y) = (cos y, cos z, cos x)
    m3 = R.Mesh([(0, 2 * math.pi, 16)] * 3, dtype=torch.float64)
    z = torch.arange(16, dtype=torch.float64) * 2 * math.pi / 16
    X3, Y3, Z3 = torch.meshgrid(z, z, z, indexing="ij")
    v3 = torch.stack([torch.sin(Z3), torch.sin(X3), torch.sin(Y3)])[None]
    c3 = torch.stack([torch.cos(Y3), torch.cos(Z3), torch.cos(X3)])[None]
    assert (m3.ifft(R.curl(m3.fft(v3), m3)).real - c3).abs().max() < 1e-12


def test_conservative_convection_closure_closed_forms():
    """The closure's ConservativeConvection term (dealiased input, no aliasing for these modes):
    1D u = sin x + cos(2x)/2 gives d_x(u^2) = 2 u u_x; 2D u = (sin x, sin y) gives
    (sin 2x + sin x cos y, cos x sin y + sin 2y)."""
    n = 32
    x = torch.arange(n, dtype=torch.float64) * 2 * math.pi / n
    u = (torch.sin(x) + 0.5 * torch.cos(2 * x)).reshape(1, 1, n)
    ux = (torch.cos(x) - torch.sin(2 * x)).reshape(1, 1, n)
    o = R.apply([(1, "cons", None)], [(0, 2 * math.pi, n)], u)
    assert (o - 2 * u * ux).abs().max() < 1e-12
    X, Y = torch.meshgrid(x, x, indexing="ij")
    v = torch.stack([torch.sin(X), torch.sin(Y)])[None]
    ex = torch.stack([torch.sin(2 * X) + torch.sin(X) * torch.cos(Y), torch.cos(X) * torch.sin(Y) + torch.sin(2 * Y)])[None]
    o2 = R.apply([(1, "cons", None)], [(0, 2 * math.pi, n)] * 2, v)
    assert (o2 - ex).abs().max() < 1e-12


# ---------------------------------------------------------------- refpath's SETDRK3 and ETDRK1 branches
@pytest.mark.parametrize("scheme", ["SETDRK3", "ETDRK1"])
def test_refpath_setdrk3_etdrk1_match_numref_f64(scheme):
    """refpath's SETDRK3 (_uncached.py:120-160, _setdrk_step.py:28-52) and ETDRK1 (_etdrk.py:30-51)
    branches in float64 against numref's independent restatement (phi-function coefficients,
    half spectrum): 1D Burgers N = 128 for 200 steps and forced Navier-Stokes 64^2 for 20 steps."""
    g = golden("g1_burgers1d.npz")
    u0 = torch.from_numpy(g["u0_f64"])
    o = R.integrate(R.burgers_terms(0.01), [(0, 1, 128)], u0, 0.01, 200, scheme=scheme)
    t = numref.integrate([(0, 1, 128)], [(0.01, "lap", None), (-1, "conv", None)], g["u0_f64"], 0.01, 200, scheme)
    assert rel(o.numpy(), t) < 1e-12
    g = golden("g2_ns64.npz")
    y = kolm_y(64, torch.float64)
    o = R.integrate(R.ns_terms(1000, y), [(0, 2 * math.pi, 64)] * 2, torch.from_numpy(g["u0_f64"]), 1e-3, 20,
                    scheme=scheme)
    yn = np.broadcast_to(2 * np.pi * np.arange(64) / 64, (1, 1, 64, 64))
    terms = [(-1, "vc", None), (1 / 1000, "lap", None), (-0.1, "unit", None), (-1, "src", 4 * np.cos(4 * yn))]
    t = numref.integrate([(0, 2 * np.pi, 64)] * 2, terms, g["u0_f64"], 1e-3, 20, scheme)
    assert rel(o.numpy(), t) < 1e-12


@pytest.mark.parametrize("scheme,order", [("SETDRK3", 3), ("ETDRK1", 1)])
def test_refpath_observed_order(scheme, order):
    """Observed order of refpath's own SETDRK3 and ETDRK1 steps on 1D viscous Burgers in float64,
    against a 400-step SETDRK4 solution: a wrong coefficient or stage formula loses the order."""
    n = 64
    x = torch.arange(n, dtype=torch.float64) / n
    u0 = (torch.sin(2 * math.pi * x) + 0.5).reshape(1, 1, n)
    terms = R.burgers_terms(0.02)
    T = 0.2
    ref = R.integrate(terms, [(0, 1, n)], u0, T / 400, 400, scheme="SETDRK4")
    errs = [float((R.integrate(terms, [(0, 1, n)], u0, T / m, m, scheme=scheme) - ref).abs().max())
            for m in (10, 20, 40)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert abs(rates[-1] - order) < 0.3, rates


@pytest.mark.parametrize("scheme", ["ETDRK1", "SETDRK1", "SETDRK3"])
def test_refpath_constant_forcing_exact(scheme):
    """u' = L u + s with a constant source: ETDRK1's and SETDRK1's step is exact,
    u(t) = e^{Lt} u0 + t phi1(Lt) s per mode, and so is SETDRK3's (its c3 + c4 + c5 = dt phi1).
    refpath in float64 against that closed form."""
    n, nu, dt, steps = 16, 0.05, 0.1, 5
    x = torch.arange(n, dtype=torch.float64) / n
    src = torch.cos(2 * math.pi * 2 * x).reshape(1, 1, n)
    u0 = torch.sin(2 * math.pi * x).reshape(1, 1, n)
    o = R.integrate([(nu, "lap", None), (1.0, "src", src)], [(0, 1, n)], u0, dt, steps, scheme=scheme)
    t = steps * dt
    lam = -nu * (2 * math.pi * 2) ** 2
    ex = math.exp(-nu * (2 * math.pi) ** 2 * t) * u0 + t * (math.exp(lam * t) - 1) / (lam * t) * src
    assert float((o - ex).abs().max()) < 1e-12


def _hermitian_part(x, dims):
    """(X(k) + conj(X(-k))) / 2 over the given FFT dims."""
    xm = torch.roll(torch.flip(x, dims=dims), shifts=[1] * len(dims), dims=dims)
    return 0.5 * (x + xm.conj())


@pytest.mark.parametrize("dims,n", [(2, 16), (3, 8)])
def test_nyquist_plane_parity_rule_is_the_hermitian_part(dims, n):
    """Terms that differentiate after a product's FFT (ConservativeConvection, the projection of
    NSPressureConvection) are not Hermitian on Nyquist planes in the reference's full spectrum.
    Their Hermitian part, which a real output keeps, equals the engine's rule: on axis d's Nyquist
    plane a product of symbols survives only with an even number of Nyquist-axis factors a_d, except
    at modes that are their own conjugate partner, which keep the full value.  Checked in float64
    on fields with content at every frequency (no dealiasing); and the reference's real output after
    10 steps of each scheme equals stepping with the Hermitian part of N."""
    torch.manual_seed(3)
    mi = [(0, 1, n)] * dims
    ax = tuple(range(-dims, 0))
    mesh = R.Mesh(mi, dtype=torch.float64)
    u = torch.randn(2, dims, *([n] * dims), dtype=torch.float64)
    uf = torch.fft.fftn(u, dim=ax)
    a = torch.meshgrid(*[2 * math.pi * torch.fft.fftfreq(n, 1 / n, dtype=torch.float64)] * dims, indexing="ij")
    idx = torch.meshgrid(*[torch.arange(n)] * dims, indexing="ij")
    ny = [i == n // 2 for i in idx]
    own = torch.ones_like(ny[0])
    for d in range(dims):
        own = own & ((idx[d] == 0) | ny[d])
    ny = [f & ~own for f in ny]
    for kind in ("cons", "nsp"):
        raw = R.build_closure(mesh, [(1, kind, None)], dealias_rate=10.0)(uf)
        # own-partner modes keep the full value (the real output takes their real part either way)
        ref = torch.where(own, raw, _hermitian_part(raw, ax))
        got = torch.zeros_like(uf)
        if kind == "cons":
            for c in range(dims):
                for i in range(dims):
                    ai = torch.where(ny[i], 0.0, a[i])
                    got[:, c] += 1j * ai * torch.fft.fftn(u[:, i] * u[:, c], dim=ax)
        else:
            conv = R.convection(uf, mesh, u)
            lap = sum(-(a[i] ** 2) for i in range(dims))
            il = torch.where(lap == 0, 1.0, 1 / lap)
            for c in range(dims):
                acc = sum(torch.where(ny[c] == ny[i], -a[c] * a[i], 0.0) * conv[:, i] for i in range(dims))
                got[:, c] = acc * il - conv[:, c]
        e = float((got - ref).abs().max() / ref.abs().max())
        assert e < 1e-13, (kind, e)
    # the reference's real output is the dynamics of the Hermitian part
    for kind in ("cons", "nsp"):
        terms = [(0.02, "lap", None), (-1 if kind == "cons" else 1, kind, None)]
        u0 = 0.3 * torch.randn(1, dims, *([n] * dims), dtype=torch.float64)
        for scheme in ("SETDRK4", "ETDRK2"):
            ref = R.integrate(terms, mi, u0, 1e-3, 10, scheme=scheme)
            orig = R.build_closure
            try:
                R.build_closure = lambda m, t, r=2 / 3: (lambda f, cl=orig(m, t, r): _hermitian_part(cl(f), ax))
                emu = R.integrate(terms, mi, u0, 1e-3, 10, scheme=scheme)
            finally:
                R.build_closure = orig
            assert R.rel_l2(emu, ref) < 1e-13, (kind, scheme)


@pytest.mark.parametrize("dims", [2, 3])
def test_refpath_forced_ns_velocity_vs_reference_golden(dims):
    """NavierStokes(100, force=ExplicitSource(f)) (pde.py:85-100): the forced core masks the
    integrator's stage tensors in place and adds the force twice (_navier_stokes.py:238-262).  The
    oracle's restatement matches the reference's float64 run (SETDRK4) to 1e-12, every scheme's
    float32 run to 1e-5, and one __call__ (golden g10)."""
    g = golden("g10_nsp_force.npz")
    tag = f"{dims}d"
    n = g[f"u0_{tag}"].shape[-1]
    mi = [(0, 1, n)] * dims
    terms = [(1, "nsp_f", torch.from_numpy(g[f"force_{tag}"])), (1 / 100, "lap", None)]
    u64 = torch.from_numpy(g[f"u0_{tag}"])
    out = R.integrate(terms, mi, u64, 2e-3, 6, scheme="SETDRK4")
    assert R.rel_l2(out, g[f"out_{tag}_setdrk4_f64"]) < 1e-12
    t32 = [(1, "nsp_f", torch.from_numpy(g[f"force_{tag}"]).float()), (1 / 100, "lap", None)]
    for scheme in ("ETDRK1", "ETDRK2", "SETDRK1", "SETDRK2", "SETDRK3", "SETDRK4"):
        out = R.integrate(t32, mi, u64.float(), 2e-3, 6, scheme=scheme)
        assert R.rel_l2(out, g[f"out_{tag}_{scheme.lower()}_f32"]) < 1e-5, scheme
    call = R.apply(terms, mi, u64, return_in_fourier=True)
    assert R.rel_l2(call, g[f"call_{tag}_f64"]) < 1e-12


def test_forced_ns_core_doubles_a_divergence_free_force():
    """At u = 0 the forced core returns grad(invlap div(-f)) + 2 f, which is 2 f for a
    divergence-free f: f = (sin 2 pi y, sin 2 pi x) on [0,1)^2."""
    n = 16
    mesh = R.Mesh([(0, 1, n)] * 2, dtype=torch.float64)
    x = torch.arange(n, dtype=torch.float64) / n
    X, Y = torch.meshgrid(x, x, indexing="ij")
    f = torch.stack([torch.sin(2 * math.pi * Y), torch.sin(2 * math.pi * X)])[None]
    ff = torch.fft.fftn(f, dim=(-2, -1))
    u = torch.zeros(1, 2, n, n, dtype=torch.complex128)
    got = R.ns_pressure_convection_forced(u, mesh, ff)
    assert float((got - 2 * ff).abs().max()) < 1e-12
