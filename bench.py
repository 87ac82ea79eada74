#!/usr/bin/env python
"""Benchmark: 2D incompressible Navier-Stokes (vorticity form) with Kolmogorov forcing on a 512^2
grid, batch 64, nu = 1e-3 (Re = 1000), SETDRK4, dt = 1e-3 -- the throughput workload of the ETD
engine, in DOF*steps/s (DOF = B * C * 512^2).

One step = one SETDRK4 step of the whole batch (4 nonlinear evaluations: 4 row + 4 column
kernels), state resident in HBM.  Prints ONE JSON line on stdout; progress goes to stderr.

Legs:
  engine       K timed steps after W warm-up steps (torch.cuda.synchronize on both sides)
  kernels      HIP-event timing of the row and column kernels on the engine's stream -> roofline
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

N, B, RE, DT = 512, 64, 1000.0, 1e-3
HBM_PEAK_GBS = 8000.0
METRIC = "DOF·steps/s, 2D NS 512² b64 vs torch path; p50/p99 latency, mixed load"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_engine(dev, w0):
    import torch

    import torchfsm_amd as tfa
    from torchfsm_amd import field, pde
    from torchfsm_amd.integrator import SETDRKIntegrator

    mesh = tfa.MeshGrid([(0, 2 * math.pi, N)] * 2, device=dev)
    y = mesh.bc_mesh_grid()[1]
    op = pde.NavierStokesVorticity(RE, field.kolm_force(y))
    op.set_integrator(SETDRKIntegrator.SETDRK4)
    op.register_mesh(tfa.FourierMesh(mesh), 1)
    plan = op._prepare(B, dev, DT)
    uh = plan.forward(w0.to(dev))
    torch.cuda.synchronize()
    return op, plan, uh, y


def time_kernels(plan, uh, reps=20):
    """Average duration (ms) of the row kernel and of the column kernel per SETDRK4 stage, from HIP
    events recorded on the stream the kernels are launched on."""
    import torch

    u = uh.clone()
    stream = torch.cuda.current_stream()
    out = {}
    plan.launch_pass(u, 1, 0)  # fields of the state
    plan.launch_pass(u, 0)
    for name, pid, stage in [("row", 0, 0), ("col_e0", 1, 7), ("col_e1", 1, 8), ("col_e2", 1, 9), ("col_e3", 1, 10)]:
        plan.launch_pass(u, pid, stage)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            plan.launch_pass(u, pid, stage)
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--torch-steps", type=int, default=20)
    ap.add_argument("--no-baselines", action="store_true", help="engine leg only (used under the profiler)")
    ap.add_argument("--tile", type=int, default=0, help="column tile (8/16); 0 = autotune the launch shape")
    ap.add_argument("--col-var", type=int, default=0)
    ap.add_argument("--row-var", type=int, default=0)
    ap.add_argument("--row-wg", type=int, default=256)
    ap.add_argument("--sub-batch", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 3 plain 256 MiB device copies (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). The batch shards across GPUs with no communication; see DESIGN.md.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from torchfsm_amd.workloads import ns_initial_vorticity

    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] copy probe: 3 x 256 MiB device-to-device copies done")
    log(f"[bench] NS {N}^2 B={B} SETDRK4 dt={DT}: building input and plan")
    w0 = ns_initial_vorticity(N, B, seed=0)
    op, plan, uh, y = build_engine(dev, w0)
    if args.tile:
        plan.set_tuning(args.tile, args.col_var, args.row_var, args.row_wg, args.sub_batch)
    else:
        from torchfsm_amd.autotune import autotune

        autotune(plan, uh, steps=6, log=log)
    log(f"[bench] launch shape (col tile, col var, row var, row wg, sub-batch) = {plan.tuning}; "
        f"warmup {args.warmup} steps")
    plan.step(uh, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(uh, args.steps)
    torch.cuda.synchronize()
    t_eng = time.perf_counter() - t0
    dof = B * 1 * N * N
    value = dof * args.steps / t_eng
    ms_step = t_eng / args.steps * 1e3
    finite = bool(torch.isfinite(plan.inverse(uh)).all().item())
    log(f"[bench] engine: {ms_step:.4f} ms/step, {value:.4e} DOF*steps/s, finite={finite}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Gaussian random vorticity, E(k)~k^-3 on 1<=|k|<=32, unit RMS",
        "config": {
            "workload": "ns2d_kolmogorov_512x512_b64_setdrk4",
            "grid": [N, N],
            "batch": B,
            "nu": 1.0 / RE,
            "dt": DT,
            "scheme": "SETDRK4",
            "dealias": "2/3",
            "forcing": "kolm_force(y): -0.1 w - 4 cos(4y)",
            "launch_shape": {"col_tile": plan.tuning[0], "col_radix_variant": plan.tuning[1],
                             "row_radix_variant": plan.tuning[2], "row_workgroup": plan.tuning[3],
                             "sub_batch": plan.tuning[4]},
            "parallelism": "single GPU, batch-parallel",
        },
    }
    if args.no_baselines:
        print(json.dumps(result), flush=True)
        return

    # ---------------------------------------------------------------- kernel roofline
    kt = time_kernels(plan, uh)
    H = N // 2 + 1
    kmax = plan_kmax = N // 3  # |k| <= floor(N/3) kept (2/3 rule)
    A = B * N * H * 8  # bytes of one half-spectrum field over the batch
    rho = (kmax + 1) / H
    row_bytes = (4 * rho + 1) * A
    col_bytes = {"col_e0": 3 * A, "col_e1": 3 * A, "col_e2": 5 * A, "col_e3": 5 * A}
    col_bytes = {k: v + 4 * rho * A for k, v in col_bytes.items()}
    t_row_step = 4 * kt["row"]
    t_col_step = sum(kt[k] for k in col_bytes)
    step_bytes = 4 * row_bytes + sum(col_bytes.values())
    if t_col_step >= t_row_step:
        dom = ("k2_col<512,tile,VC> (column: fwd x-FFT + SETDRK4 combine + 4 inverse x-FFTs)",
               sum(col_bytes.values()) / 4, t_col_step / 4)
    else:
        dom = ("k2_row<512,4,VC> (row: 2 packed C2R + product + R2C)", row_bytes, kt["row"])
    achieved = dom[1] / (dom[2] * 1e-3) / 1e9
    result["roofline"] = {
        "bound": "hbm",
        "kernel": dom[0],
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "kernel_ms": {k: round(v, 5) for k, v in kt.items()},
        "step_bytes_model": step_bytes,
        "step_frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    log(f"[bench] kernels (ms): {kt}; dominant {dom[0]} {achieved:.0f} GB/s")

    # ---------------------------------------------------------------- baselines (oracle restatement)
    from oracle import refpath

    terms_cpu = refpath.ns_terms(RE, y.cpu())
    threads = min(16, os.cpu_count() or 1)
    torch.set_num_threads(threads)
    log(f"[bench] cpu baseline: refpath float32 on {threads} threads, B={B}, {args.cpu_steps} steps")
    mesh_info = [(0, 2 * math.pi, N)] * 2
    mesh_c = refpath.Mesh(mesh_info, device="cpu")
    L = refpath.build_linear(mesh_c, terms_cpu, 1)
    sch = refpath.Scheme("SETDRK4", DT, L, refpath.build_closure(mesh_c, terms_cpu))
    uc = mesh_c.fft(w0)
    uc = sch.step(uc)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        uc = sch.step(uc)
    t_cpu = time.perf_counter() - t0
    cpu_val = dof * args.cpu_steps / t_cpu
    result["cpu_baseline"] = {
        "value": cpu_val,
        "unit": "DOF·steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle refpath (torch CPU, MKL FFT={torch.backends.mkl.is_available()}), full workload "
                  f"B={B} 512^2, 1 warm-up + {args.cpu_steps} timed SETDRK4 steps ({t_cpu:.1f} s)",
    }
    log(f"[bench] cpu baseline {cpu_val:.3e} DOF*steps/s ({t_cpu:.1f} s)")

    log(f"[bench] torch-GPU baseline: refpath on {dev}, {args.torch_steps} steps")
    mesh_g = refpath.Mesh(mesh_info, device=dev)
    terms_g = refpath.ns_terms(RE, y)
    Lg = refpath.build_linear(mesh_g, terms_g, 1)
    schg = refpath.Scheme("SETDRK4", DT, Lg, refpath.build_closure(mesh_g, terms_g))
    ug = mesh_g.fft(w0.to(dev))
    for _ in range(2):
        ug = schg.step(ug)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.torch_steps):
        ug = schg.step(ug)
    torch.cuda.synchronize()
    t_tg = time.perf_counter() - t0
    tg_val = dof * args.torch_steps / t_tg
    result["torch_gpu_baseline"] = {
        "value": tg_val,
        "unit": "DOF·steps/s",
        "ms_per_step": t_tg / args.torch_steps * 1e3,
        "kind": "port",
        "sample": f"oracle refpath (full complex torch.fft, reference op order) on the GPU, B={B}, "
                  f"2 warm-up + {args.torch_steps} timed steps",
        "engine_speedup": value / tg_val,
    }
    log(f"[bench] torch-GPU baseline {tg_val:.3e} DOF*steps/s; engine speedup {value / tg_val:.2f}x")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
