#!/usr/bin/env python
"""Benchmark of the ETD engine.  Prints ONE JSON line on stdout; progress goes to stderr.

Workloads (--workload):
  ns2d       (default) 2D incompressible Navier-Stokes, vorticity form, Kolmogorov forcing,
             512^2 grid, batch 64, nu = 1e-3 (Re = 1000), SETDRK4, dt = 1e-3.  One step = one
             SETDRK4 step of the whole batch (4 nonlinear evaluations: 4 row + 4 column kernels).
  burgers3d  3D vector Burgers(0.01), 128^3 grid, 3 channels, batch 8, ETDRK2, dt = 5e-3.  One
             step = 2 evaluations, each a z-row, y-forward, x-column and y-inverse kernel.  At
             these settings the reference itself diverges for 4 of the 8 samples (advective CFL
             above the ETDRK2 limit); the engine tracks it step for step (tests), so the output is
             not finite after ~20 steps.  Throughput does not depend on the values.
The metric is DOF*steps/s with DOF = B * C * prod(N), state resident in HBM.

Legs:
  engine       K timed steps after W warm-up steps (torch.cuda.synchronize on both sides)
  kernels      HIP-event timing of each kernel of the step on the engine's stream, launched with
               the same sub-batch shape as the step -> roofline of the dominant kernel (runs
               before the engine leg, so the GPU is out of idle when the engine leg starts)
  differentiation  (ns2d) the Burgers 128^2 adjoint workload: integrate() + backward() against
               the reference's method (torch autograd through the oracle restatement)
  cpu_baseline the oracle's restatement of the reference path on the host cores (bounded sample)
  torch_gpu    the same restatement on the GPU (torch.fft / rocFFT), the "torch path"
"""

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "DOF·steps/s, 2D NS 512² b64 vs torch path; p50/p99 latency, mixed load"

# stage codes of the fused kernels (engine.hpp Stage)
ST_PRO, ST_R2_E0, ST_R2_E1, ST_R4_E0, ST_R4_E1, ST_R4_E2, ST_R4_E3 = 0, 2, 3, 7, 8, 9, 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class NS2D:
    name = "ns2d_kolmogorov_512x512_b64_setdrk4"
    N, B, C, RE, DT, SCHEME = 512, 64, 1, 1000.0, 1e-3, "SETDRK4"
    default_steps = 1000
    cpu_steps = 15  # about 10 s of the oracle on the host cores
    traffic_file = "k2_col_ns512_b64_traffic.txt"

    def __init__(self, dev):
        import torch

        import torchfsm_amd as tfa
        from torchfsm_amd import field, pde
        from torchfsm_amd.integrator import SETDRKIntegrator
        from torchfsm_amd.workloads import ns_initial_vorticity

        self.dev = dev
        self.u0 = ns_initial_vorticity(self.N, self.B, seed=0)
        mesh = tfa.MeshGrid([(0, 2 * math.pi, self.N)] * 2, device=dev)
        self.y = mesh.bc_mesh_grid()[1]
        op = pde.NavierStokesVorticity(self.RE, field.kolm_force(self.y))
        op.set_integrator(SETDRKIntegrator.SETDRK4)
        op.register_mesh(tfa.FourierMesh(mesh), 1)
        self.plan = op._prepare(self.B, dev, self.DT)
        self.uh = self.plan.forward(self.u0.to(dev))
        torch.cuda.synchronize()
        self.dof = self.B * self.C * self.N * self.N

    def config(self):
        return {"workload": self.name, "grid": [self.N, self.N], "batch": self.B, "nu": 1.0 / self.RE,
                "dt": self.DT, "scheme": self.SCHEME, "dealias": "2/3",
                "forcing": "kolm_force(y): -0.1 w - 4 cos(4y)"}

    data = "synthetic: seeded Gaussian random vorticity, E(k)~k^-3 on 1<=|k|<=32, unit RMS"

    def kernels(self, nb):
        """(name, pass id, stage, algorithmic bytes per launch, launches per step).  A = one
        half-spectrum field of the launch's nb samples; rho = kept fraction of ky (2/3 rule)."""
        H = self.N // 2 + 1
        A = nb * self.N * H * 8
        rho = (self.N // 3 + 1) / H
        per = self.B // nb
        row = ("k2_row (512-point rows: 2 packed C2R, product u.grad w, R2C)", 0, ST_R4_E0, (4 * rho + 1) * A, 4 * per)
        cols = [(f"k2_col (512-point columns: fwd FFT, SETDRK4 combine, 4 inverse FFTs) e{i}", 1, st,
                 (r + 4 * rho) * A, per)
                for i, (st, r) in enumerate([(ST_R4_E0, 3), (ST_R4_E1, 3), (ST_R4_E2, 5), (ST_R4_E3, 3)])]
        return [row] + cols

    def prime(self):
        """Fill the product/field buffers so single passes see a valid state."""
        self.plan.launch_pass(self.uh.clone(), 1, ST_PRO)

    prime_pass = (1, ST_PRO)

    def step_sequence(self, nb):
        """The launches of one step, in order: (kernels() name, pass id, stage)."""
        ks = self.kernels(nb)
        row, cols = ks[0][0], [k[0] for k in ks[1:]]
        seq = []
        for i, st in enumerate((ST_R4_E0, ST_R4_E1, ST_R4_E2, ST_R4_E3)):
            seq += [(row, 0, st), (cols[i], 1, st)]
        return seq

    def cpu_baseline(self, args):
        import torch

        from oracle import refpath

        terms = refpath.ns_terms(self.RE, self.y.cpu())
        mesh_info = [(0, 2 * math.pi, self.N)] * 2
        mesh_c = refpath.Mesh(mesh_info, device="cpu")
        L = refpath.build_linear(mesh_c, terms, 1)
        sch = refpath.Scheme("SETDRK4", self.DT, L, refpath.build_closure(mesh_c, terms))
        uc = sch.step(mesh_c.fft(self.u0))
        t0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            uc = sch.step(uc)
        t = time.perf_counter() - t0
        return self.dof * args.cpu_steps / t, (
            f"oracle refpath (torch CPU, MKL FFT={torch.backends.mkl.is_available()}), full workload "
            f"B={self.B} 512^2, 1 warm-up + {args.cpu_steps} timed SETDRK4 steps ({t:.1f} s)")

    def torch_gpu(self, steps):
        from oracle import refpath

        mesh_info = [(0, 2 * math.pi, self.N)] * 2
        mesh_g = refpath.Mesh(mesh_info, device=self.dev)
        terms = refpath.ns_terms(self.RE, self.y)
        sch = refpath.Scheme("SETDRK4", self.DT, refpath.build_linear(mesh_g, terms, 1),
                             refpath.build_closure(mesh_g, terms))
        return sch, mesh_g.fft(self.u0.to(self.dev)), self.dof


class Burgers3D:
    name = "burgers3d_128cubed_c3_b8_etdrk2"
    N, B, C, NU, DT, SCHEME = 128, 8, 3, 0.01, 5e-3, "ETDRK2"
    default_steps = 200
    cpu_steps = 100  # one sample: about 10 s of the oracle on the host cores
    traffic_file = "k3_burgers128_b8_traffic.txt"
    data = "synthetic: per (b, c) 8 seeded cosines, integer |k| in [1, 4], a ~ U(-1,1), phase ~ U(0, 2pi)"

    def __init__(self, dev):
        import torch

        import torchfsm_amd as tfa
        from torchfsm_amd import pde
        from torchfsm_amd.integrator import ETDRKIntegrator
        from torchfsm_amd.workloads import fourier_modes_field

        self.dev = dev
        self.u0 = fourier_modes_field(self.N, 3, 3, self.B, seed=0)
        self.mesh_info = [(0, 1, self.N)] * 3
        op = pde.Burgers(self.NU)
        op.set_integrator(ETDRKIntegrator.ETDRK2)
        op.register_mesh(tfa.FourierMesh(self.mesh_info, device=dev), 3)
        self.plan = op._prepare(self.B, dev, self.DT)
        self.uh = self.plan.forward(self.u0.to(dev))
        torch.cuda.synchronize()
        self.dof = self.B * self.C * self.N ** 3

    def config(self):
        return {"workload": self.name, "grid": [self.N] * 3, "channels": 3, "batch": self.B, "nu": self.NU,
                "dt": self.DT, "scheme": self.SCHEME, "dealias": "2/3"}

    def kernels(self, nb):
        """A1 = one half-spectrum channel of the launch's nb samples; rz = kept kz fraction,
        ryz = kept (ky, kz) fraction.  Fields G2 hold kept kz only, G1 kept (ky, kz) only."""
        N = self.N
        H = N // 2 + 1
        A1 = nb * N * N * H * 8
        k = N // 3
        rz = (k + 1) / H
        ryz = (2 * k + 1) / N * rz
        per = self.B // nb
        return [
            ("k3_row (z rows: 12 C2R, products, 3 R2C)", 4, ST_R2_E0, (9 * rz + 3) * A1, 2 * per),
            ("k_cols_c2c (y forward of the products)", 5, ST_R2_E0, 6 * A1, 2 * per),
            ("k3_col (x fwd + combine + 6 x inverse) e0", 6, ST_R2_E0, (9 + 6 * ryz) * A1, per),
            ("k3_col (x fwd + combine + 6 x inverse) e1", 6, ST_R2_E1, (12 + 6 * ryz) * A1, per),
            ("k3_yinv (6 -> 9 fields, y inverse)", 7, ST_R2_E0, (6 * ryz + 9 * rz) * A1, 2 * per),
        ]

    def prime(self):
        self.plan.launch_pass(self.uh.clone(), 3, ST_PRO)

    prime_pass = (3, ST_PRO)

    def step_sequence(self, nb):
        ks = {k[0]: k for k in self.kernels(nb)}
        names = list(ks)
        row, yfwd, col0, col1, yinv = names
        seq = []
        for st, col in ((ST_R2_E0, col0), (ST_R2_E1, col1)):
            seq += [(row, 4, st), (yfwd, 5, st), (col, 6, st), (yinv, 7, st)]
        return seq

    def cpu_baseline(self, args):
        import torch

        from oracle import refpath

        # bounded sample: one sample of the batch, 1 warm-up + cpu_steps timed steps
        terms = refpath.burgers_terms(self.NU)
        mesh_c = refpath.Mesh(self.mesh_info, device="cpu")
        sch = refpath.Scheme(self.SCHEME, self.DT, refpath.build_linear(mesh_c, terms, 3),
                             refpath.build_closure(mesh_c, terms))
        uc = sch.step(mesh_c.fft(self.u0[:1]))
        t0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            uc = sch.step(uc)
        t = time.perf_counter() - t0
        dof1 = self.C * self.N ** 3
        return dof1 * args.cpu_steps / t, (
            f"oracle refpath (torch CPU, MKL FFT={torch.backends.mkl.is_available()}), 1 of the {self.B} "
            f"samples at 128^3 x 3 channels, 1 warm-up + {args.cpu_steps} timed {self.SCHEME} steps ({t:.1f} s)")

    def torch_gpu(self, steps):
        from oracle import refpath

        mesh_g = refpath.Mesh(self.mesh_info, device=self.dev)
        terms = refpath.burgers_terms(self.NU)
        sch = refpath.Scheme(self.SCHEME, self.DT, refpath.build_linear(mesh_g, terms, 3),
                             refpath.build_closure(mesh_g, terms))
        return sch, mesh_g.fft(self.u0.to(self.dev)), self.dof


WORKLOADS = {"ns2d": NS2D, "burgers3d": Burgers3D}

SERVE_SIZES = {"ks1d": (512, 1024, 2048, 4096), "ns2d": (64, 128, 256, 512), "burgers3d": (64, 128)}
SERVE_MIX = ("KS 1D N in {512..4096}, L ~ U[32pi, 128pi], dt 0.1 SETDRK4 | NS 2D N in {64..512}, nu ~ logU(1e-3, "
             "5e-2), dt 1e-3 SETDRK4 | 3D Burgers 64^3 and 128^3, nu ~ logU(1e-2, 5e-2), dt 2e-3 SETDRK2; "
             "steps ~ U{100..2000}")
SERVE_P99_BOUND_S = 2.0  # the p99 latency the rate sweep is held to


def serve_stream(dev, reqs):
    """Serve `reqs` through the continuous-batching server (classes created before the clock starts)."""
    import torch

    from torchfsm_amd import serving

    srv = serving.Server(dev)
    t0 = time.perf_counter()
    srv.prewarm(sorted({(r.kind, r.n) for r in reqs}))
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = srv.run(reqs)
    t_run = time.perf_counter() - t0
    work = sum(serving.SPECS[r.kind][1] * r.n ** serving.SPECS[r.kind][0] * r.steps for r in reqs)
    lat = serving.latency_summary(res)
    out = {
        "requests": len(res), "makespan_s": t_run, "setup_s": t_setup,
        "dof_steps_per_s": work / t_run, "requests_per_s": len(res) / t_run,
        "p50_latency_ms": lat["p50_ms"], "p99_latency_ms": lat["p99_ms"], "max_latency_ms": lat["max_ms"],
        "all_finite": all(bool(torch.isfinite(r.u).all()) for r in res),
        "by_class": serving.latency_by_class(res),
    }
    srv.close()
    return out, work


def plain_solves(dev, reqs, warm):
    """The same requests solved one at a time.  cold: integrate() on a new operator per request (plan
    and coefficient setup included, as a per-request solve pays it).  warm: one single-slot plan per
    shape class, kept across requests; each request re-targets it (tfsm_plan_set_sample), transforms
    its field, steps and transforms back."""
    import torch

    from torchfsm_amd import field, pde, serving
    from torchfsm_amd.integrator import SETDRKIntegrator
    from torchfsm_amd.mesh import MeshGrid

    classes = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        dims, ch, dt, scheme, length = serving.SPECS[r.kind]
        if warm:
            c = classes.get((r.kind, r.n))
            if c is None:
                c = classes[(r.kind, r.n)] = serving.ShapeClass(r.kind, r.n, 1, dev)
            c.plan.set_sample(0, serving.linear_coefs(r.kind, r.param), float(r.param) if r.kind == "ks1d" else 0.0)
            c.plan.forward(r.u0.to(dev), out=c.uh)
            c.plan.step(c.uh, r.steps)
            c.plan.inverse(c.uh).cpu()
            continue
        if r.kind == "ks1d":
            op, mesh = pde.KuramotoSivashinsky(), MeshGrid([(0, r.param, r.n)], device=dev)
        elif r.kind == "ns2d":
            mesh = MeshGrid([(0, length, r.n)] * 2, device=dev)
            op = pde.NavierStokesVorticity(r.param, field.kolm_force(mesh.bc_mesh_grid()[1]))
        else:
            mesh = MeshGrid([(0, length, r.n)] * 3, device=dev)
            op = pde.Burgers(r.param)
        op.set_integrator(getattr(SETDRKIntegrator, scheme))
        op.integrate(r.u0.to(dev), mesh=mesh, dt=dt, step=r.steps).cpu()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    for c in classes.values():
        c.close()
    return t


def serving_leg(dev, n_requests, rate, seed=0, rates=(), rate_requests=60):
    """The mixed request stream through the continuous-batching server; the same requests solved one
    by one (warm and cold per-request baselines); and a sweep of lighter arrival rates for the
    highest rate whose p99 latency stays within SERVE_P99_BOUND_S."""
    from torchfsm_amd import serving

    reqs = serving.synthetic_requests(n_requests, rate, seed=seed, sizes=SERVE_SIZES, steps=(100, 2000))
    out, work = serve_stream(dev, reqs)
    out = {"arrival_rate_per_s": rate, **out, "mix": SERVE_MIX}
    log(f"[bench] serving: {out['requests']} requests in {out['makespan_s']:.2f} s, "
        f"{out['dof_steps_per_s']:.3e} DOF*steps/s, p50 {out['p50_latency_ms']:.1f} ms, "
        f"p99 {out['p99_latency_ms']:.1f} ms")
    for warm in (True, False):
        t = plain_solves(dev, reqs, warm)
        name = "plain_per_request_warm" if warm else "plain_per_request_cold"
        out[name] = {"requests": len(reqs), "seconds": t, "dof_steps_per_s": work / t,
                     "throughput_vs_plain": out["dof_steps_per_s"] / (work / t)}
        log(f"[bench] {name}: {t:.2f} s for the same {len(reqs)} requests; served "
            f"{out[name]['throughput_vs_plain']:.2f}x")
    sweep = []
    for i, r in enumerate(rates):
        lreqs = serving.synthetic_requests(rate_requests, r, seed=seed + 1 + i, sizes=SERVE_SIZES, steps=(100, 2000))
        lo, _ = serve_stream(dev, lreqs)
        sweep.append({"arrival_rate_per_s": r, "requests": lo["requests"], "requests_per_s": lo["requests_per_s"],
                      "p50_latency_ms": lo["p50_latency_ms"], "p99_latency_ms": lo["p99_latency_ms"],
                      "by_class": lo["by_class"]})
        log(f"[bench] serving at {r}/s: p50 {lo['p50_latency_ms']:.1f} ms, p99 {lo['p99_latency_ms']:.1f} ms")
    if sweep:
        met = [x["arrival_rate_per_s"] for x in sweep if x["p99_latency_ms"] <= SERVE_P99_BOUND_S * 1e3]
        out["rate_sweep"] = sweep
        out["p99_bound_ms"] = SERVE_P99_BOUND_S * 1e3
        out["max_rate_within_p99_bound"] = max(met) if met else None
    return out


def differentiation_leg(dev, reps=5):
    """The differentiation workload: 2D viscous Burgers 128^2, C=2, B=16, nu=0.01, 100 ETDRK2 steps
    of dt 0.01, loss = MSE against a seeded target, gradients w.r.t. u0 and nu through the
    hand-written adjoint with a checkpoint every 10 steps.  Timed: integrate() alone, and
    integrate() + loss.backward(); beside it the reference's method (torch autograd through the
    oracle restatement's full-complex torch.fft step, float32 on the same GPU, u0 gradient)."""
    import torch

    import torchfsm_amd as tfa
    from torchfsm_amd import pde
    from torchfsm_amd.integrator import ETDRKIntegrator
    from torchfsm_amd.workloads import fourier_modes_field

    n, c, b, steps, dt = 128, 2, 16, 100, 0.01
    u0 = (0.1 * fourier_modes_field(n, c, 2, b, seed=21)).to(dev)
    target = (0.1 * fourier_modes_field(n, c, 2, b, seed=22)).to(dev)
    mesh = tfa.MeshGrid([(0, 1, n)] * 2, device=dev)
    nu = torch.tensor(0.01, device=dev, requires_grad=True)
    op = pde.Burgers(nu)
    op.set_integrator(ETDRKIntegrator.ETDRK2)
    op.adjoint_checkpoint_every = 10
    op_fwd = pde.Burgers(0.01)  # the same operator with a plain coefficient: integrate() alone
    op_fwd.set_integrator(ETDRKIntegrator.ETDRK2)

    def run(grad):
        u = u0.clone().requires_grad_(grad)
        if grad:
            nu.grad = None
        res = (op if grad else op_fwd).integrate(u, mesh=mesh, dt=dt, step=steps)
        if grad:
            # the operator's coefficient (nu times the Laplacian's unit coefficient) is one graph
            # node shared by every run: keep it for the next backward
            ((res - target) ** 2).mean().backward(retain_graph=True)
        return res

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2] * 1e3

    with torch.no_grad():
        fwd_ms = timed(lambda: run(False))
    fb_ms = timed(lambda: run(True))

    from oracle import refpath

    def ref():
        u = u0.clone().requires_grad_(True)
        out = refpath.integrate(refpath.burgers_terms(0.01), [(0, 1, n)] * 2, u, dt, steps, scheme="ETDRK2")
        ((out - target) ** 2).mean().backward()

    ref_ms = timed(ref)
    out = {"workload": f"burgers2d_{n}x{n}_c{c}_b{b}_etdrk2_{steps}steps_ckpt10",
           "forward_ms": fwd_ms, "forward_backward_ms": fb_ms, "backward_ms": fb_ms - fwd_ms,
           "backward_ms_per_step": (fb_ms - fwd_ms) / steps,
           "backward_over_forward": (fb_ms - fwd_ms) / fwd_ms,
           "reference_autograd_ms": ref_ms, "speedup_vs_reference_autograd": ref_ms / fb_ms,
           "timing": f"median of {reps} wall-clock runs (device synchronised) after one warm-up",
           "reference_method": "torch autograd through the oracle restatement of the reference step "
                               "(full complex torch.fft, float32, same GPU), u0 gradient"}
    log(f"[bench] differentiation: forward {fwd_ms:.2f} ms, forward+backward {fb_ms:.2f} ms "
        f"(backward {fb_ms - fwd_ms:.2f} ms); reference autograd {ref_ms:.2f} ms")
    return out


CPU_SHARE = 16  # host CPUs allotted to one GPU's jobs on the machines this runs on


def host_cpu():
    """CPU model, os.cpu_count(), the CPUs this process may run on, and the CPU FFT backend."""
    import torch

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "os_cpu_count": os.cpu_count(), "usable_cpus": usable,
            "mkl": bool(torch.backends.mkl.is_available())}


def time_in_step(plan, uh, prime, seq, steps=12, skip=2):
    """Average duration (ms) of each kernel of the step, timed inside the step's own schedule:
    every sub-batch runs `steps` steps launched pass by pass in the step's order, each kernel
    recording a pair of HIP events at its own start and end (hipExtLaunchKernelGGL through
    tfsm_launch_pass_timed, so launch gaps are not counted).  The sub-batches run one after the
    other on torch's current stream, so each duration is the kernel's own (as a profiler that
    serialises kernels sees it); the engine leg overlaps two sub-batches where the plan uses two
    streams, which the step-level fraction reflects.  The first `skip` steps of every sub-batch are
    not counted."""
    import torch

    B, sb = plan.batch, plan.tuning[4]
    nsub = 1 if not 0 < sb < B else (B + sb - 1) // sb
    main = torch.cuda.current_stream()
    streams = [main]
    u = uh.clone()
    torch.cuda.synchronize()
    fork = torch.cuda.Event()
    fork.record(main)
    evs = []  # per sub-batch: [steps][len(seq) + 1] events
    for k in range(nsub):
        s = streams[k % len(streams)]
        s.wait_event(fork)
        with torch.cuda.stream(s):
            plan.launch_pass(u, *prime, sub=k)
            ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in seq]
                  for _ in range(steps)]
            for pair in (p for row in ev for p in row):  # torch creates an event at its first record
                for e in pair:
                    e.record(s)
            for i in range(steps):
                for j, (_, pid, st) in enumerate(seq):
                    # the kernel itself records the pair at its start and end (tfsm_launch_pass_timed)
                    plan.launch_pass(u, pid, st, sub=k, events=ev[i][j])
        evs.append(ev)
    torch.cuda.synchronize()
    tot, cnt = {}, {}
    for ev in evs:
        for i in range(skip, steps):
            for j, (name, _, _) in enumerate(seq):
                tot[name] = tot.get(name, 0.0) + ev[i][j][0].elapsed_time(ev[i][j][1])
                cnt[name] = cnt.get(name, 0) + 1
    return {n: tot[n] / cnt[n] for n in tot}


def time_kernel(plan, uh, pass_id, stage, reps=20):
    """Average duration (ms) of one pass, from HIP events recorded on the stream the engine
    launches on (torch's current stream)."""
    import torch

    stream = torch.cuda.current_stream()
    plan.launch_pass(uh, pass_id, stage)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.launch_pass(uh, pass_id, stage)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def read_traffic(fname, tuning):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain-copy calibration), or None when
    the summary is absent or was taken at another launch shape."""
    path = os.path.join(HERE, "profiles", fname)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("launch_shape") != ",".join(map(str, tuning)):
        return None
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="ns2d")
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0 = the workload's default)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=0, help="timed oracle steps on the host (0 = the workload's default)")
    ap.add_argument("--torch-steps", type=int, default=20)
    ap.add_argument("--no-baselines", action="store_true", help="engine and kernel legs only")
    ap.add_argument("--tile", type=int, default=0, help="column tile (8/16); 0 = the plan's default launch shape")
    ap.add_argument("--autotune", action="store_true", help="search the launch shape first (torchfsm_amd.autotune)")
    ap.add_argument("--col-var", type=int, default=0)
    ap.add_argument("--row-var", type=int, default=0)
    ap.add_argument("--row-wg", type=int, default=256)
    ap.add_argument("--sub-batch", type=int, default=0)
    ap.add_argument("--streams", type=int, default=1, help="1, or 2 to overlap alternate sub-batches")
    ap.add_argument("--serve-requests", type=int, default=240, help="requests of the serving leg (0 = skip it)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--serve-rates", default="5,10,20,40",
                    help="lighter Poisson rates (1/s, comma-separated) swept for the p99 bound; empty: none")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 3 plain 256 MiB device copies (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-differentiation", action="store_true",
                    help="skip the differentiation leg (Burgers 128^2 adjoint vs reference autograd)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine has no multi-GPU path (--gpus {args.gpus} requested). "
              f"The batch would shard across GPUs with no communication.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] copy probe: 3 x 256 MiB device-to-device copies done")
    W = WORKLOADS[args.workload]
    steps = args.steps or W.default_steps
    args.cpu_steps = args.cpu_steps or W.cpu_steps
    log(f"[bench] {W.name}: building input and plan")
    w = W(dev)
    plan, uh = w.plan, w.uh
    if args.tile:
        plan.set_tuning(args.tile, args.col_var, args.row_var, args.row_wg, args.sub_batch)
        plan.set_streams(args.streams)
    elif args.autotune:
        from torchfsm_amd.autotune import autotune

        autotune(plan, uh, steps=6, log=log)
    log(f"[bench] launch shape (col tile, col var, row var, row wg, sub-batch) = {plan.tuning}, "
        f"streams {plan.streams}; warmup {args.warmup} steps")
    # the kernel leg (below) runs first: a dozen steps of timed launches, which also bring the GPU
    # out of idle before the engine leg's W warmup and K timed steps
    sb = plan.tuning[4]
    nb = sb if 0 < sb < w.B else w.B
    ks = w.kernels(nb)
    kt = time_in_step(plan, uh, w.prime_pass, w.step_sequence(nb))
    plan.step(uh, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(uh, steps)
    torch.cuda.synchronize()
    t_eng = time.perf_counter() - t0
    value = w.dof * steps / t_eng
    ms_step = t_eng / steps * 1e3
    finite = bool(torch.isfinite(plan.inverse(uh)).all().item())
    log(f"[bench] engine: {ms_step:.4f} ms/step, {value:.4e} DOF*steps/s, finite={finite}")

    cfg = w.config()
    cfg["launch_shape"] = {"col_tile": plan.tuning[0], "col_radix_variant": plan.tuning[1],
                           "row_radix_variant": plan.tuning[2], "row_workgroup": plan.tuning[3],
                           "sub_batch": plan.tuning[4], "streams": plan.streams}
    cfg["parallelism"] = "single GPU, batch-parallel"
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF·steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": w.data,
        "config": cfg,
    }

    # ---------------------------------------------------------------- kernel roofline
    # dominant kernel family = the largest share of the step's kernel time
    fam = {}
    for name, pid, stage, nbytes, per_step in ks:
        f = name.rsplit(" e", 1)[0] if name[-2:-1] == "e" and name[-1].isdigit() else name
        t, b, n = fam.get(f, (0.0, 0.0, 0))
        fam[f] = (t + kt[name] * per_step, b + nbytes * per_step, n + per_step)
    dom = max(fam, key=lambda f: fam[f][0])
    t_dom, b_dom, n_dom = fam[dom]
    avg_ms = t_dom / n_dom
    avg_bytes = b_dom / n_dom
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    step_bytes = sum(nbytes * per_step for _, _, _, nbytes, per_step in ks)
    traffic = read_traffic(W.traffic_file, plan.tuning)
    result["roofline"] = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "algorithmic_bytes_per_launch": avg_bytes,
        "avg_launch_ms": avg_ms,
        "launch_samples": nb,
        "kernel_ms": {k: round(v, 5) for k, v in kt.items()},
        "step_bytes_model": step_bytes,
        "step_frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "streams": plan.streams,
        "timing": ("in-step: HIP events recorded by each kernel at its start and end, every sub-batch's steps in "
                   "order, the sub-batches one after the other on one stream (each kernel's own duration)" +
                   ("; the engine leg overlaps two sub-batches on two streams (step_frac is the step-level "
                    "fraction)" if plan.streams == 2 else "")),
    }
    log(f"[bench] kernels (ms): {kt}; dominant {dom}: {avg_bytes / 1e6:.1f} MB per launch in {avg_ms:.4f} ms "
        f"= {achieved:.0f} GB/s; traffic {traffic}")
    if args.no_baselines:
        print(json.dumps(result), flush=True)
        return

    # ---------------------------------------------------------------- baselines (oracle restatement)
    host = host_cpu()
    # the host cores this process may use, capped at the 16 CPUs allotted per GPU
    threads = min(CPU_SHARE, host["usable_cpus"])
    torch.set_num_threads(threads)
    log(f"[bench] cpu baseline: refpath float32 on {threads} threads ({host['model']}, os.cpu_count() "
        f"{host['os_cpu_count']}), {args.cpu_steps} steps")
    cpu_val, sample = w.cpu_baseline(args)
    result["cpu_baseline"] = {"value": cpu_val, "unit": "DOF·steps/s", "cores": threads, "kind": "port",
                              "sample": sample, "host": host,
                              "threads_note": f"threads = min({CPU_SHARE}, usable CPUs): {CPU_SHARE} host "
                                              f"CPUs are allotted per GPU"}
    log(f"[bench] cpu baseline {cpu_val:.3e} DOF*steps/s")

    log(f"[bench] torch-GPU baseline: refpath on {dev}, {args.torch_steps} steps")
    sch, ug, dof = w.torch_gpu(args.torch_steps)
    for _ in range(2):
        ug = sch.step(ug)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.torch_steps):
        ug = sch.step(ug)
    torch.cuda.synchronize()
    t_tg = time.perf_counter() - t0
    tg_val = dof * args.torch_steps / t_tg
    result["torch_gpu_baseline"] = {
        "value": tg_val,
        "unit": "DOF·steps/s",
        "ms_per_step": t_tg / args.torch_steps * 1e3,
        "kind": "port",
        "sample": f"oracle refpath (full complex torch.fft, reference op order) on the GPU, whole batch, "
                  f"2 warm-up + {args.torch_steps} timed steps",
        "engine_speedup": value / tg_val,
    }
    log(f"[bench] torch-GPU baseline {tg_val:.3e} DOF*steps/s; engine speedup {value / tg_val:.2f}x")
    if args.workload == "ns2d":
        del sch, ug, w, plan, uh
        torch.cuda.empty_cache()
        if not args.no_differentiation:
            result["differentiation"] = differentiation_leg(dev)
    if args.serve_requests > 0 and args.workload == "ns2d":
        torch.cuda.empty_cache()
        log(f"[bench] serving: {args.serve_requests} requests at {args.serve_rate}/s")
        rates = [float(x) for x in args.serve_rates.split(",") if x.strip()] if args.serve_rates else []
        result["serving"] = serving_leg(dev, args.serve_requests, args.serve_rate, rates=rates)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
